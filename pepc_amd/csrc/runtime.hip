// runtime.hip -- context, device memory and the C-ABI of include/pepc_gpu.h.
//
// One context owns the device buffers of one particle system: particle
// arrays sized for n (grown, never shrunk, so a timestep loop allocates only
// on its first step), node arrays sized node_factor * n, and the small
// device scalar block the kernels report through.  grow and traverse are
// sequences of kernel launches on the context's stream with no host round
// trip; the host reads the scalars once, in pepc_gpu_sync().
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pepc_gpu.h"
#include "kernels.h"
#include "lattice_table.h"

using namespace pepc;

namespace {

struct Ctx {
  pepc_gpu_config cfg;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  int64_t n = 0, ncap = 0, mcap = 0, io_cap = 0;
  int64_t node_min = 0;   // lower bound of the node pool (pepc_gpu_reserve_nodes)
  bool grown = false, walked = false;
  int walk_blocks = 0;
  int nxcd = 1;   // XCDs of the device: one walk work queue each

  // particles
  double *in[4] = {nullptr, nullptr, nullptr, nullptr};   // host-path input copies
  unsigned long long *k0 = nullptr, *k1 = nullptr, *oldkeys = nullptr;
  uint32_t *v0 = nullptr, *v1 = nullptr;
  double *xs = nullptr, *ys = nullptr, *zs = nullptr, *qs = nullptr;
  double4* prec4 = nullptr;         // {x, y, z, q} records, input order
  uint8_t *lo = nullptr, *hi = nullptr;
  uint32_t *cnt = nullptr, *gmark = nullptr, *dflag = nullptr, *dlist = nullptr;
  long long* kd = nullptr;
  uint32_t *gfirst = nullptr, *gcount = nullptr, *gstat = nullptr, *queue = nullptr;
  uint32_t *np2p = nullptr, *nm2p = nullptr;
  double *res[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // host-path outputs
  // nodes
  uint32_t* tab = nullptr;          // BFS numbering table, bfs_table_words(ncap) words
  uint8_t* ntype = nullptr;         // node kinds (0 leaf, 1 twig)
  uint8_t* kidtwig = nullptr;       // per twig: which children are twigs
  uint32_t *fc_rank = nullptr,
           *node_rank = nullptr, *nfirst = nullptr, *nparent = nullptr;
  NodeLink* link = nullptr;
  NodeGeom* geom = nullptr;
  double* mom64 = nullptr;
  NodeMoments<float>* wm32 = nullptr;
  NodeMoments<double>* wm64 = nullptr;
  // scratch
  uint32_t *hist = nullptr, *bsum = nullptr;
  int64_t hist_cap = 0, bsum_cap = 0;
  double* partial = nullptr;
  DevBox* box = nullptr;
  DevScalars* sc = nullptr;
  DevScalars* sc_host = nullptr;    // pinned
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  TreeBuffers tb;                   // buffers of the last grow
  // batched systems
  int64_t nsys = 1, sys_cap = 0;
  std::vector<int64_t> hoff;        // offsets on the device (soff), host copy
  uint32_t* sys = nullptr;          // system per particle (input and sorted order)
  int64_t* soff = nullptr;
  uint32_t* sysroot = nullptr;
  // on-device timestep
  bool has_state = false, capturing = false;
  double* st[11] = {};              // x y z vx vy vz q m ax ay az
  double* tmp[4] = {};              // gathered vx vy vz m
  long long *label = nullptr, *tlabel = nullptr;
  double* ehist = nullptr;          // energies: record x system x {pot, kin}
  int64_t ehist_cap = 0, nrec = 0;
  uint32_t* ecount = nullptr;
  int64_t* echunk = nullptr;        // chunk boundaries of the energy reduction
  int* esyschunk = nullptr;         // first chunk of each system
  double* epartial = nullptr;
  int nechunk = 0;
  hipStream_t cstream = nullptr;    // capture stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double graph_dt = 0.0;
  // adjoint
  double* nadj = nullptr;
  uint32_t *parent = nullptr, *leafof = nullptr, *aqueue = nullptr;
  int64_t adj_mcap = 0, adj_ncap = 0;
  // frozen interaction lists
  uint32_t *lcount = nullptr, *loff = nullptr, *lent = nullptr;
  unsigned long long* ltotal = nullptr;
  int64_t list_ncap = 0, lent_cap = 0, frozen_entries = 0;
  bool frozen = false;
  // periodic boundaries
  bool periodic = false, far_field = false;
  pepc_gpu_periodic per;
  LatticeGeom lgeom;
  int nbox = 1;
  double* vbox = nullptr;           // 3 * nbox lattice vectors, central box last
  double* fact = nullptr;           // factorials 0..80
  double2 *ml = nullptr, *mstar = nullptr, *lstar = nullptr, *mu = nullptr, *lpart = nullptr;
  double *ldpart = nullptr, *fict = nullptr, *linv = nullptr, *ldip = nullptr, *lcoef = nullptr;
  int* lfpt = nullptr;              // merged fictitious-charge points (LatticeGeom::fpt)
  double* lfpos = nullptr;          // their positions and charges (LatticeGeom::fpos)
  double2* lom = nullptr;
  // sinks (traversal of other particles)
  int64_t scap = 0, sk_m = 0;
  int sk_mode = 0;
  double *sk_x = nullptr, *sk_y = nullptr, *sk_z = nullptr;
  double* sk_res[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* sk_io[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  unsigned long long *sk_k0 = nullptr, *sk_k1 = nullptr;
  uint32_t *sk_v0 = nullptr, *sk_v1 = nullptr, *sk_gfirst = nullptr, *sk_gcount = nullptr;
  uint32_t *sk_p2p = nullptr, *sk_m2p = nullptr, *sk_hist = nullptr, *sk_bsum = nullptr;
  // where the last sink traversal left its counts and order (the context's
  // sink buffers, or the no-allocation scratch of the device)
  const uint32_t *sk_last_p2p = nullptr, *sk_last_m2p = nullptr, *sk_last_perm = nullptr;
};

// Device buffers of one sink traversal.
struct SinkBufs {
  double *x, *y, *z;              // sorted sink coordinates (grouped mode)
  double* res[5];                 // results in sorted order (grouped mode)
  unsigned long long *k0, *k1;    // sort keys (ping-pong)
  uint32_t *v0, *v1;              // sort values (ping-pong); v0 = sorted -> input index
  uint32_t *gfirst, *gcount;      // walk groups (grouped mode)
  uint32_t *p2p, *m2p;            // per-sink interaction counts
  uint32_t *hist, *bsum;          // radix histogram and scan blocks
};

// ---------------------------------------------------------------- scratch
// Device memory of pepc_gpu_traverse_sinks_noalloc: one static array per
// device, carved into the SinkBufs of up to kNoAllocSinks sinks.  Host code
// never uses the array's name as an address: each device's copy is looked
// up once (hipGetSymbolAddress at context creation, with the device current)
// and that pointer is what the kernels receive.  All calls on one device
// share its copy, so calls that overlap on one device race.  Every entry a
// traversal reads is first written by the same traversal (keys, values,
// histograms, groups, sorted coordinates, results, counts), so nothing needs
// clearing between calls.
constexpr int64_t kNoAllocSinks = 262144;
constexpr size_t scratch_pad(size_t b) { return (b + 255) / 256 * 256; }
constexpr int64_t kNoAllocSortBlocks = (kNoAllocSinks + 2047) / 2048;   // >= sort_blocks(kNoAllocSinks)
constexpr int64_t kNoAllocHist = 256 * kNoAllocSortBlocks;
constexpr int64_t kNoAllocBsum = 512;   // >= scan_blocks(kNoAllocHist) + 64 and the 256 digit totals
constexpr size_t kScratchD = scratch_pad(sizeof(double) * kNoAllocSinks);
constexpr size_t kScratchK = scratch_pad(sizeof(unsigned long long) * kNoAllocSinks);
constexpr size_t kScratchU = scratch_pad(sizeof(uint32_t) * kNoAllocSinks);
constexpr size_t kScratchBytes = 8 * kScratchD + 2 * kScratchK + 6 * kScratchU +
                                 scratch_pad(sizeof(uint32_t) * kNoAllocHist) +
                                 scratch_pad(sizeof(uint32_t) * kNoAllocBsum);   // 26.2 MB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};           // device address of each device's scratch
hipError_t g_scratch_err[kMaxDevices] = {};      // lookup failure, if any

// Once per device, with that device current.
void scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    g_scratch_err[dev] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    (void)hipGetLastError();
    return;
  }
  g_scratch_ptr[dev] = ptr;
  g_scratch_err[dev] = hipSuccess;
}

SinkBufs scratch_sink_bufs(void* base) {
  unsigned char* p = static_cast<unsigned char*>(base);
  SinkBufs b;
  auto take = [&p](size_t bytes) {
    unsigned char* q = p;
    p += bytes;
    return q;
  };
  b.x = reinterpret_cast<double*>(take(kScratchD));
  b.y = reinterpret_cast<double*>(take(kScratchD));
  b.z = reinterpret_cast<double*>(take(kScratchD));
  for (int k = 0; k < 5; ++k) b.res[k] = reinterpret_cast<double*>(take(kScratchD));
  b.k0 = reinterpret_cast<unsigned long long*>(take(kScratchK));
  b.k1 = reinterpret_cast<unsigned long long*>(take(kScratchK));
  b.v0 = reinterpret_cast<uint32_t*>(take(kScratchU));
  b.v1 = reinterpret_cast<uint32_t*>(take(kScratchU));
  b.gfirst = reinterpret_cast<uint32_t*>(take(kScratchU));
  b.gcount = reinterpret_cast<uint32_t*>(take(kScratchU));
  b.p2p = reinterpret_cast<uint32_t*>(take(kScratchU));
  b.m2p = reinterpret_cast<uint32_t*>(take(kScratchU));
  b.hist = reinterpret_cast<uint32_t*>(take(scratch_pad(sizeof(uint32_t) * kNoAllocHist)));
  b.bsum = reinterpret_cast<uint32_t*>(take(scratch_pad(sizeof(uint32_t) * kNoAllocBsum)));
  return b;
}

int fail(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
  return fail(c, PEPC_GPU_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHECK(c, expr)                                  \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return hip_fail((c), _e, #expr); \
  } while (0)

template <typename T>
int dalloc(Ctx* c, T** p, int64_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count <= 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, PEPC_GPU_ENOMEM,
                std::string("hipMalloc of ") + std::to_string(sizeof(T) * (size_t)count) +
                    " bytes failed: " + hipGetErrorString(e));
  }
  return 0;
}

#define ALLOC(c, p, count)                 \
  do {                                     \
    int _r = dalloc((c), &(p), (count));   \
    if (_r) return _r;                     \
  } while (0)

void free_all(Ctx* c) {
  void* ptrs[] = {c->in[0], c->in[1], c->in[2], c->in[3], c->k0, c->k1, c->oldkeys, c->v0, c->v1,
                  c->xs, c->ys, c->zs, c->qs, c->lo, c->hi, c->cnt, c->gmark, c->dflag, c->dlist,
                  c->kd, c->gfirst, c->gcount, c->gstat, c->queue, c->np2p, c->nm2p, c->res[0], c->res[1],
                  c->res[2], c->res[3], c->res[4], c->tab, c->ntype, c->kidtwig,
                  c->fc_rank, c->node_rank, c->nfirst, c->link, c->geom,
                  c->mom64, c->wm32, c->wm64, c->hist, c->bsum, c->partial, c->box, c->sc,
                  c->sys, c->soff, c->sysroot, c->label, c->tlabel, c->ehist, c->ecount,
                  c->nadj, c->parent, c->leafof, c->aqueue, c->lcount, c->loff, c->lent,
                  c->ltotal, c->nparent, c->echunk, c->esyschunk, c->epartial, c->prec4,
                  c->vbox, c->fact, c->ml, c->mstar, c->lstar, c->mu, c->lpart, c->ldpart,
                  c->fict, c->linv, c->ldip, c->lom, c->lcoef, c->lfpt, c->lfpos, c->sk_x, c->sk_y, c->sk_z,
                  c->sk_res[0], c->sk_res[1], c->sk_res[2], c->sk_res[3], c->sk_res[4],
                  c->sk_io[0], c->sk_io[1], c->sk_io[2], c->sk_io[3], c->sk_io[4], c->sk_io[5],
                  c->sk_io[6], c->sk_io[7], c->sk_k0, c->sk_k1, c->sk_v0, c->sk_v1, c->sk_gfirst,
                  c->sk_gcount, c->sk_p2p, c->sk_m2p, c->sk_hist, c->sk_bsum};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (double* p : c->st)
    if (p) (void)hipFree(p);
  for (double* p : c->tmp)
    if (p) (void)hipFree(p);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

void drop_graph(Ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

// Node arrays for mc nodes, and the radix histogram and scan blocks sized
// for the larger of the particle and node arrays.  Drops the current tree.
int alloc_nodes(Ctx* c, int64_t mc) {
  // node indices (links, parents) are 32-bit, and the walk's stack entries
  // keep the node's kind in the top bit: the pool must stay below 2^31
  if ((double)mc >= 2147483648.0)
    return fail(c, PEPC_GPU_EINVAL, "node pool exceeds 31-bit node indices");
  const int64_t nc = c->ncap;
  ALLOC(c, c->fc_rank, mc + 1); ALLOC(c, c->ntype, mc); ALLOC(c, c->kidtwig, mc); ALLOC(c, c->node_rank, mc); ALLOC(c, c->nfirst, mc);
  ALLOC(c, c->nparent, mc);
  ALLOC(c, c->link, mc); ALLOC(c, c->geom, mc); ALLOC(c, c->mom64, mc * 16);
  ALLOC(c, c->wm32, mc); ALLOC(c, c->wm64, mc);
  // radix histogram: 256 digits x sort blocks of the particle arrays; scan
  // blocks for the largest array scanned (particle-sized)
  const int64_t hc = 256 * sort_blocks(nc);
  ALLOC(c, c->hist, hc);
  int64_t bc = scan_blocks(hc > nc ? hc : nc) + 64;
  if (bc < 512) bc = 512;   // also holds the 256 digit totals of a radix pass
  ALLOC(c, c->bsum, bc);
  c->mcap = mc;
  c->grown = false;
  c->walked = false;
  c->frozen = false;
  drop_graph(c);
  return 0;
}

int ensure_particles(Ctx* c, int64_t n) {
  if (n <= c->ncap) return 0;
  const int64_t nc = n + n / 8 + 64;
  int64_t mc = (int64_t)(c->cfg.node_factor * (double)nc) + 64;
  if (c->cfg.node_factor * (double)nc + 64.0 >= 2147483648.0)
    return fail(c, PEPC_GPU_EINVAL, "node pool of node_factor * n nodes exceeds 31-bit node indices");
  if (mc < c->node_min) mc = c->node_min;
  if (mc < nc) mc = nc;   // every particle is a leaf (and the grow borrows node arrays of n entries)
  ALLOC(c, c->k0, nc); ALLOC(c, c->k1, nc); ALLOC(c, c->oldkeys, nc);
  ALLOC(c, c->v0, nc); ALLOC(c, c->v1, nc);
  ALLOC(c, c->xs, nc); ALLOC(c, c->ys, nc); ALLOC(c, c->zs, nc); ALLOC(c, c->qs, nc);
  ALLOC(c, c->prec4, nc);
  ALLOC(c, c->lo, nc); ALLOC(c, c->hi, nc);
  ALLOC(c, c->cnt, nc); ALLOC(c, c->gmark, nc); ALLOC(c, c->dflag, nc); ALLOC(c, c->dlist, nc);
  ALLOC(c, c->kd, nc);
  ALLOC(c, c->gfirst, nc); ALLOC(c, c->gcount, nc); ALLOC(c, c->gstat, 8 * nc);
  ALLOC(c, c->sys, nc);
  ALLOC(c, c->np2p, nc); ALLOC(c, c->nm2p, nc);
  ALLOC(c, c->tab, bfs_table_words(nc));
  c->ncap = nc;
  int r = alloc_nodes(c, mc);
  if (r) return r;
  c->io_cap = 0;        // staging and state arrays follow the particle capacity
  c->has_state = false;
  return 0;
}

int ensure_systems(Ctx* c, int64_t nsys) {
  if (nsys <= c->sys_cap) return 0;
  ALLOC(c, c->box, nsys);
  ALLOC(c, c->soff, nsys + 1);
  ALLOC(c, c->sysroot, nsys);
  c->sys_cap = nsys;
  c->hoff.clear();
  drop_graph(c);
  return 0;
}

// Uploads the system ranges when they differ from the ones on the device.
int set_offsets(Ctx* c, int64_t nsys, const int64_t* off) {
  int r = ensure_systems(c, nsys);
  if (r) return r;
  if ((int64_t)c->hoff.size() == nsys + 1 && std::equal(c->hoff.begin(), c->hoff.end(), off))
    return 0;
  if (c->capturing) return fail(c, PEPC_GPU_EINVAL, "system ranges changed during graph capture");
  c->hoff.assign(off, off + nsys + 1);
  HIPCHECK(c, hipMemcpyAsync(c->soff, c->hoff.data(), sizeof(int64_t) * (nsys + 1),
                             hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  return 0;
}

int ensure_host_io(Ctx* c, int64_t n) {
  // device staging buffers of the host-pointer entry points (also the force
  // outputs of the on-device timestep)
  if (n <= c->io_cap && c->res[0]) return 0;
  for (int k = 0; k < 4; ++k) ALLOC(c, c->in[k], c->ncap);
  for (int k = 0; k < 5; ++k) ALLOC(c, c->res[k], c->ncap);
  c->io_cap = c->ncap;
  return 0;
}

__global__ void reset_scalars_kernel(DevScalars* sc) {
  if (threadIdx.x == 0) {
    DevScalars z;
    memset(&z, 0, sizeof(z));
    *sc = z;
  }
}

int check_flags(Ctx* c, uint32_t f) {
  if (f & kFlagBadBox) return fail(c, PEPC_GPU_EBOX, "non-finite coordinates or degenerate bounding box");
  if (f & kFlagNodeOverflow)
    return fail(c, PEPC_GPU_EOVERFLOW, "node pool overflow: raise node_factor (cf. np_mult)");
  if (f & kFlagDupKeys)
    return fail(c, PEPC_GPU_EKEYS, "duplicate keys could not be resolved (key space exhausted)");
  if (f & kFlagStackOverflow) return fail(c, PEPC_GPU_EWALK, "walk stack overflow");
  return 0;
}

Ctx* as_ctx(void* p) { return reinterpret_cast<Ctx*>(p); }

int record(Ctx* c, int k) {
  if (!c->capturing) HIPCHECK(c, hipEventRecord(c->ev[k], c->stream));
  return 0;
}

#define RECORD(c, k)              \
  do {                            \
    int _r = record((c), (k));    \
    if (_r) return _r;            \
  } while (0)

// nsys == 1: one system (no offsets needed).  nsys > 1: particles of system s
// occupy [off[s], off[s+1]) on input and, sorted by key, on output.
int grow_device_impl(Ctx* c, int64_t n, const double* x, const double* y, const double* z,
                     const double* q, int64_t nsys, const int64_t* off) {
  if (n < 1) return fail(c, PEPC_GPU_EINVAL, "grow: need at least one particle");
  if (n > (int64_t)0x7fffffff) return fail(c, PEPC_GPU_EINVAL, "grow: n exceeds 2^31-1");
  if (nsys < 1 || nsys > 65536) return fail(c, PEPC_GPU_EINVAL, "grow: nsys must be in [1, 65536]");
  if (nsys > 1) {
    if (!off) return fail(c, PEPC_GPU_EINVAL, "grow: null system offsets");
    if (off[0] != 0 || off[nsys] != n)
      return fail(c, PEPC_GPU_EINVAL, "grow: offsets must start at 0 and end at n");
    for (int64_t k = 0; k < nsys; ++k)
      if (off[k + 1] < off[k]) return fail(c, PEPC_GPU_EINVAL, "grow: offsets must not decrease");
  }
  int r = ensure_particles(c, n);
  if (r) return r;
  r = ensure_systems(c, nsys);
  if (r) return r;
  if (nsys > 1) {
    r = set_offsets(c, nsys, off);
    if (r) return r;
  }
  const uint32_t* sys = nsys > 1 ? c->sys : nullptr;
  hipStream_t s = c->stream;
  RECORD(c, 0);
  reset_scalars_kernel<<<1, 64, 0, s>>>(c->sc);
  if (nsys > 1)
    launch_box_batch(s, (int)nsys, c->soff, x, y, z, c->cfg.force_cubic_domain, c->box, c->sys, c->sc);
  else
    launch_box(s, n, x, y, z, c->cfg.force_cubic_domain, c->partial, c->box, c->sc);
  launch_keys(s, n, x, y, z, q, c->box, sys, c->k0, c->v0, c->prec4);
  launch_sort_keys(s, c->k0, c->v0, c->k1, c->v1, n, c->hist, c->bsum, sys, nsys);
  launch_dup_detect(s, n, c->k0, sys, c->sc);
  // equal keys into the reference presort's order (the sort's ping-pong
  // buffers, the group scratch arrays and the duplicate fix-up's arrays,
  // all written afresh before their own use, are free here)
  launch_tie_order(s, n, nsys, c->soff, c->k0, c->v0, c->k1, c->v1, c->cnt,
                   reinterpret_cast<uint16_t*>(c->gmark), c->dflag,
                   reinterpret_cast<unsigned long long*>(c->kd), c->sc);
  launch_gather_rec(s, n, c->v0, c->prec4, c->xs, c->ys, c->zs, c->qs);
  launch_dupfix(s, n, c->k0, c->xs, c->ys, c->zs, c->box, sys, nsys > 1 ? c->soff : nullptr, c->kd,
                c->oldkeys, c->dflag, c->dlist, c->bsum, c->sc);
  TreeBuffers t;
  t.n = n;
  t.cap = c->mcap;
  t.keys = c->k0;
  t.x = c->xs; t.y = c->ys; t.z = c->zs; t.q = c->qs;
  t.lo = c->lo; t.hi = c->hi; t.cnt = c->cnt;
  t.tab = c->tab; t.fc_rank = c->fc_rank; t.ntype = c->ntype; t.kidtwig = c->kidtwig;
  t.node_rank = c->node_rank; t.nfirst = c->nfirst; t.link = c->link; t.geom = c->geom;
  t.parent = c->nparent;
  t.mom64 = c->mom64; t.wm32 = c->wm32; t.wm64 = c->wm64;
  t.gmark = c->gmark; t.gfirst = c->gfirst; t.gcount = c->gcount;
  t.hist = c->hist; t.bsum = c->bsum; t.box = c->box; t.sc = c->sc;
  t.prec = c->cfg.precision == PEPC_GPU_FP64 ? kWalkFP64 : kWalkFP32Kahan;
  t.mac_select = c->cfg.mac_select;
  t.sys = sys;
  t.soff = nsys > 1 ? c->soff : nullptr;
  t.sysroot = c->sysroot;
  t.nsys = nsys;
  c->tb = t;
  launch_tree_build(s, t, c->cfg.group_size, c->cfg.group_pack);
  HIPCHECK(c, hipGetLastError());
  RECORD(c, 1);
  c->n = n;
  c->nsys = nsys;
  c->frozen = false;
  c->grown = true;
  c->walked = false;
  return 0;
}

int traverse_device_impl(Ctx* c, double* ex, double* ey, double* ez, double* pot, double* work) {
  if (!c->grown) return fail(c, PEPC_GPU_EINVAL, "traverse: no tree (call grow first)");
  if (!ex || !ey || !ez || !pot || !work) return fail(c, PEPC_GPU_EINVAL, "traverse: null output");
  hipStream_t s = c->stream;
  WalkArgs a;
  a.n = c->n;
  a.x = c->xs; a.y = c->ys; a.z = c->zs; a.q = c->qs;
  a.geom = c->geom; a.link = c->link; a.kidtwig = c->kidtwig;
  a.wm32 = c->wm32; a.wm64 = c->wm64;
  a.gfirst = c->gfirst; a.gcount = c->gcount; a.sc = c->sc;
  a.theta2 = c->cfg.theta2; a.eps2 = c->cfg.eps2; a.sign = c->cfg.sign;
  a.ex = ex; a.ey = ey; a.ez = ez; a.pot = pot; a.work = work;
  a.np2p = c->np2p; a.nm2p = c->nm2p; a.gstat = c->gstat; a.queue = c->queue; a.nqueue = c->nxcd;
  a.sys = c->nsys > 1 ? c->sys : nullptr;
  a.sysroot = c->sysroot;
  if (c->periodic && c->nsys > 1)
    return fail(c, PEPC_GPU_EINVAL, "traverse: periodic boundaries need a single system");
  a.nbox = c->periodic ? c->nbox : 1;
  a.vbox = c->periodic ? c->vbox : nullptr;
  launch_walk_reset(s, a);
  RECORD(c, 2);
  launch_walk(s, a, c->cfg.precision, c->walk_blocks, c->cfg.list_tile);
  if (c->periodic && c->far_field)
    launch_lattice_far(s, c->n, c->xs, c->ys, c->zs, c->qs, c->lgeom, c->ml, c->fact, c->linv,
                       c->cfg.sign, c->lpart, c->ldpart, c->lom, c->ldip, c->mu,
                       c->fict, c->lcoef, ex, ey, ez, pot);
  RECORD(c, 3);
  launch_walk_stats(s, a);
  HIPCHECK(c, hipGetLastError());
  c->walked = true;
  return 0;
}

int sync_impl(Ctx* c, pepc_gpu_stats* st) {
  HIPCHECK(c, hipMemcpyAsync(c->sc_host, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  const DevScalars& h = *c->sc_host;
  if (st) {
    memset(st, 0, sizeof(*st));
    st->n = c->n;
    st->nnodes = h.nnodes;
    st->ntwig = h.ntwig;
    st->nleaf = (int64_t)h.nnodes - (int64_t)h.ntwig;
    st->ngroups = h.ngroups;
    st->dup_fixed = h.dup_count;
    st->interactions = h.interactions;
    st->mac_evaluations = h.mac_evals;
    st->p2p = h.p2p;
    st->m2p = h.m2p;
    st->amb_nodes = h.amb_nodes;
    st->walk_steps = h.walk_steps;
    st->flags = h.flags;
    int ml = 0;
    for (int l = 0; l <= kMaxLevel; ++l) {
      st->level_offset[l] = h.level_offset[l];
      if (h.level_offset[l + 1] > h.level_offset[l]) ml = l;
    }
    st->level_offset[kMaxLevel + 1] = h.level_offset[kMaxLevel + 1];
    st->maxlevel = ml;
  }
  return check_flags(c, h.flags);
}

int ensure_sinks(Ctx* c, int64_t m) {
  if (m <= c->scap) return 0;
  const int64_t mc = m + m / 8 + 64;
  ALLOC(c, c->sk_x, mc); ALLOC(c, c->sk_y, mc); ALLOC(c, c->sk_z, mc);
  for (int k = 0; k < 5; ++k) ALLOC(c, c->sk_res[k], mc);
  for (int k = 0; k < 8; ++k) ALLOC(c, c->sk_io[k], mc);
  ALLOC(c, c->sk_k0, mc); ALLOC(c, c->sk_k1, mc);
  ALLOC(c, c->sk_v0, mc); ALLOC(c, c->sk_v1, mc);
  ALLOC(c, c->sk_gfirst, mc); ALLOC(c, c->sk_gcount, mc);
  ALLOC(c, c->sk_p2p, mc); ALLOC(c, c->sk_m2p, mc);
  const int64_t hc = 256 * sort_blocks(mc);
  ALLOC(c, c->sk_hist, hc);
  int64_t bc = scan_blocks(hc) + 64;
  if (bc < 512) bc = 512;
  ALLOC(c, c->sk_bsum, bc);
  c->scap = mc;
  return 0;
}

// Traversal of m sinks (device arrays) through the grown tree; outputs in
// the sinks' order.  mode 0: one sink per walk group, in the given order;
// mode 1: sinks sorted along the Hilbert curve of the source box, runs of
// group_size per group, results scattered back.
SinkBufs ctx_sink_bufs(Ctx* c) {
  SinkBufs b;
  b.x = c->sk_x; b.y = c->sk_y; b.z = c->sk_z;
  for (int k = 0; k < 5; ++k) b.res[k] = c->sk_res[k];
  b.k0 = c->sk_k0; b.k1 = c->sk_k1; b.v0 = c->sk_v0; b.v1 = c->sk_v1;
  b.gfirst = c->sk_gfirst; b.gcount = c->sk_gcount; b.p2p = c->sk_p2p; b.m2p = c->sk_m2p;
  b.hist = c->sk_hist; b.bsum = c->sk_bsum;
  return b;
}

int traverse_sinks_checks(Ctx* c, int64_t m, const double* x, const double* y, const double* z,
                          int mode, double* ex, double* ey, double* ez, double* pot, double* work) {
  if (!c->grown) return fail(c, PEPC_GPU_EINVAL, "traverse_sinks: no tree (call grow first)");
  if (c->nsys > 1) return fail(c, PEPC_GPU_EINVAL, "traverse_sinks: needs a single grown system");
  if (m < 0 || m > (int64_t)0x7fffffff) return fail(c, PEPC_GPU_EINVAL, "traverse_sinks: bad sink count");
  if (mode != PEPC_GPU_SINKS_PER_PARTICLE && mode != PEPC_GPU_SINKS_GROUPED)
    return fail(c, PEPC_GPU_EINVAL, "traverse_sinks: unknown mode");
  if (m > 0 && (!x || !y || !z || !ex || !ey || !ez || !pot || !work))
    return fail(c, PEPC_GPU_EINVAL, "traverse_sinks: null argument");
  return 0;
}

// The traversal proper on given sink buffers (enough for m sinks).
int traverse_sinks_run(Ctx* c, const SinkBufs& b, int64_t m, const double* x, const double* y,
                       const double* z, int mode, double* ex, double* ey, double* ez, double* pot,
                       double* work);

int traverse_sinks_impl(Ctx* c, int64_t m, const double* x, const double* y, const double* z,
                        int mode, double* ex, double* ey, double* ez, double* pot, double* work) {
  int r = traverse_sinks_checks(c, m, x, y, z, mode, ex, ey, ez, pot, work);
  if (r) return r;
  if (m == 0) return 0;
  r = ensure_sinks(c, m);
  if (r) return r;
  return traverse_sinks_run(c, ctx_sink_bufs(c), m, x, y, z, mode, ex, ey, ez, pot, work);
}

int traverse_sinks_run(Ctx* c, const SinkBufs& b, int64_t m, const double* x, const double* y,
                       const double* z, int mode, double* ex, double* ey, double* ez, double* pot,
                       double* work) {
  hipStream_t s = c->stream;
  WalkArgs a;
  a.n = c->n;
  a.x = c->xs; a.y = c->ys; a.z = c->zs; a.q = c->qs;
  a.geom = c->geom; a.link = c->link; a.kidtwig = c->kidtwig;
  a.wm32 = c->wm32; a.wm64 = c->wm64;
  a.sc = c->sc;
  a.theta2 = c->cfg.theta2; a.eps2 = c->cfg.eps2; a.sign = c->cfg.sign;
  a.np2p = b.p2p; a.nm2p = b.m2p; a.gstat = nullptr; a.queue = c->queue; a.nqueue = c->nxcd;
  a.sys = nullptr; a.sysroot = c->sysroot;
  a.nbox = c->periodic ? c->nbox : 1;
  a.vbox = c->periodic ? c->vbox : nullptr;
  const double *tx = x, *ty = y, *tz = z;
  double* out[5] = {ex, ey, ez, pot, work};
  if (mode == PEPC_GPU_SINKS_GROUPED) {
    launch_sink_keys(s, m, x, y, z, c->box, b.k0, b.v0);
    launch_sort_keys(s, b.k0, b.v0, b.k1, b.v1, m, b.hist, b.bsum, nullptr, 1);
    // the fourth gathered array is a scratch copy of x (gather4 moves four)
    launch_gather4(s, m, b.v0, x, y, z, x, b.x, b.y, b.z, b.res[4]);
    const int G = c->cfg.group_size;
    launch_sink_groups(s, m, G, b.gfirst, b.gcount);
    a.gfirst = b.gfirst; a.gcount = b.gcount;
    a.nsink_groups = (uint32_t)((m + G - 1) / G);
    tx = b.x; ty = b.y; tz = b.z;
    for (int k = 0; k < 5; ++k) out[k] = b.res[k];
  } else {
    a.gfirst = nullptr; a.gcount = nullptr;
    a.nsink_groups = (uint32_t)m;
  }
  a.tx = tx; a.ty = ty; a.tz = tz;
  a.ex = out[0]; a.ey = out[1]; a.ez = out[2]; a.pot = out[3]; a.work = out[4];
  launch_queue_reset(s, c->queue);
  launch_walk(s, a, c->cfg.precision, c->walk_blocks, c->cfg.list_tile);
  if (c->periodic && c->far_field) {
    launch_lattice_sources(s, c->n, c->xs, c->ys, c->zs, c->qs, c->lgeom, c->ml, c->fact, c->lpart,
                           c->ldpart, c->lom, c->ldip, c->mu, c->fict, c->lcoef);
    launch_lattice_field(s, m, tx, ty, tz, c->lgeom, c->lcoef, c->linv, c->fict, c->cfg.sign,
                         out[0], out[1], out[2], out[3]);
  }
  if (mode == PEPC_GPU_SINKS_GROUPED) {
    double* dst[5] = {ex, ey, ez, pot, work};
    for (int k = 0; k < 5; ++k) launch_scatter_back(s, m, b.v0, b.res[k], dst[k]);
  }
  HIPCHECK(c, hipGetLastError());
  c->sk_m = m;
  c->sk_mode = mode;
  c->sk_last_p2p = b.p2p;
  c->sk_last_m2p = b.m2p;
  c->sk_last_perm = b.v0;
  return 0;
}

}  // namespace

extern "C" {

void pepc_gpu_config_default(pepc_gpu_config* cfg) {
  memset(cfg, 0, sizeof(*cfg));
  cfg->theta2 = 0.36;
  cfg->eps2 = 0.0;
  cfg->sign = 1.0;
  cfg->node_factor = 2.5;
  cfg->mac_select = 0;
  cfg->group_size = 16;
  cfg->precision = PEPC_GPU_FP32_KAHAN;
  cfg->force_cubic_domain = 0;
  cfg->walk_blocks = 0;
  cfg->device = 0;
}

int pepc_gpu_create(void** out, const pepc_gpu_config* cfg) {
  if (!out || !cfg) return PEPC_GPU_EINVAL;
  *out = nullptr;
  if (cfg->group_size < 1 || cfg->group_size > 64 ||
      cfg->group_pack < 0 || cfg->group_pack > 64 ||
      !(cfg->theta2 >= 0.0) || !(cfg->eps2 >= 0.0) || !(cfg->node_factor >= 1.0) ||
      (cfg->precision != PEPC_GPU_FP32_KAHAN && cfg->precision != PEPC_GPU_FP64) ||
      !walk_list_tile_ok(cfg->precision == PEPC_GPU_FP64 ? kWalkFP64 : kWalkFP32Kahan,
                         cfg->list_tile) ||
      !(cfg->sign == 1.0 || cfg->sign == -1.0))
    return PEPC_GPU_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PEPC_GPU_EHIP;
  if (cfg->device < 0 || cfg->device >= ndev) return PEPC_GPU_EINVAL;
  if (hipSetDevice(cfg->device) != hipSuccess) return PEPC_GPU_EHIP;
  Ctx* c = new Ctx();
  c->cfg = *cfg;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PEPC_GPU_EHIP;
  }
  c->own_stream = true;
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return PEPC_GPU_EHIP;
    }
  if (dalloc(c, &c->box, 1) || dalloc(c, &c->sc, 1) || dalloc(c, &c->partial, 1024 * 6) ||
      dalloc(c, &c->queue, kQueueWords) ||
      hipHostMalloc((void**)&c->sc_host, sizeof(DevScalars), hipHostMallocDefault) != hipSuccess) {
    free_all(c);
    delete c;
    return PEPC_GPU_ENOMEM;
  }
  {
    int nx = 1;
    if (hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, cfg->device) != hipSuccess) {
      (void)hipGetLastError();   // attribute unknown to this runtime: one queue
      nx = 1;
    }
    c->nxcd = nx < 1 ? 1 : (nx > kMaxQueues ? kMaxQueues : nx);
  }
  c->walk_blocks = cfg->walk_blocks > 0
                       ? cfg->walk_blocks
                       : walk_default_blocks(cfg->precision == PEPC_GPU_FP64 ? kWalkFP64 : kWalkFP32Kahan,
                                             cfg->list_tile);
  scratch_setup(cfg->device);   // the no-allocation entry point's memory on this device
  *out = c;
  return 0;
}

int pepc_gpu_destroy(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  if (c->sc_host) (void)hipHostFree(c->sc_host);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int pepc_gpu_last_error(const void* p, char* buf, size_t len) {
  const Ctx* c = reinterpret_cast<const Ctx*>(p);
  if (!buf || len == 0) return PEPC_GPU_EINVAL;
  const std::string m = c ? c->err : std::string("no context");
  snprintf(buf, len, "%s", m.c_str());
  return 0;
}

int pepc_gpu_set_stream(void* p, void* stream) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (c->own_stream && c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  c->stream = reinterpret_cast<hipStream_t>(stream);
  c->own_stream = false;
  return 0;
}

int pepc_gpu_grow_device(void* p, int64_t n, const double* x, const double* y, const double* z,
                         const double* q) {
  Ctx* c = as_ctx(p);
  if (!c || !x || !y || !z || !q) return fail(c, PEPC_GPU_EINVAL, "grow: null argument");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  return grow_device_impl(c, n, x, y, z, q, 1, nullptr);
}

int pepc_gpu_grow_batch_device(void* p, int64_t nsys, const int64_t* offsets, const double* x,
                               const double* y, const double* z, const double* q) {
  Ctx* c = as_ctx(p);
  if (!c || !x || !y || !z || !q || !offsets) return fail(c, PEPC_GPU_EINVAL, "grow_batch: null argument");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  if (nsys < 1 || nsys > 65536) return fail(c, PEPC_GPU_EINVAL, "grow_batch: nsys must be in [1, 65536]");
  return grow_device_impl(c, offsets[nsys], x, y, z, q, nsys, offsets);
}

int pepc_gpu_traverse_device(void* p, double* ex, double* ey, double* ez, double* pot,
                             double* work) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  return traverse_device_impl(c, ex, ey, ez, pot, work);
}

int pepc_gpu_sync(void* p, pepc_gpu_stats* st) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  return sync_impl(c, st);
}

int pepc_gpu_reserve_nodes(void* p, int64_t nodes) {
  Ctx* c = as_ctx(p);
  if (!c || nodes < 0) return fail(c, PEPC_GPU_EINVAL, "reserve_nodes: bad argument");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  c->node_min = nodes;
  if (c->ncap > 0 && nodes > c->mcap) return alloc_nodes(c, nodes);
  return 0;
}

int pepc_gpu_grow(void* p, int64_t n, const double* x, const double* y, const double* z,
                  const double* q, uint64_t* keys_out, int64_t* perm_out, double* xs, double* ys,
                  double* zs) {
  Ctx* c = as_ctx(p);
  if (!c || !x || !y || !z || !q || n < 1) return fail(c, PEPC_GPU_EINVAL, "grow: bad argument");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = ensure_particles(c, n);
  if (r) return r;
  r = ensure_host_io(c, n);
  if (r) return r;
  const double* src[4] = {x, y, z, q};
  for (int k = 0; k < 4; ++k)
    HIPCHECK(c, hipMemcpyAsync(c->in[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  r = grow_device_impl(c, n, c->in[0], c->in[1], c->in[2], c->in[3], 1, nullptr);
  if (r) return r;
  r = sync_impl(c, nullptr);
  if (r == PEPC_GPU_EOVERFLOW) {
    // the build counted its nodes before it stopped: enlarge the pool to
    // that count and grow again (the inputs are still staged)
    const int64_t need = (int64_t)c->sc_host->nnodes;
    r = alloc_nodes(c, need + need / 16 + 64);
    if (r) return r;
    r = grow_device_impl(c, n, c->in[0], c->in[1], c->in[2], c->in[3], 1, nullptr);
    if (r) return r;
    r = sync_impl(c, nullptr);
  }
  if (r) return r;
  if (keys_out)
    HIPCHECK(c, hipMemcpy(keys_out, c->k0, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
  if (perm_out) {
    uint32_t* tmp = new uint32_t[n];
    hipError_t e = hipMemcpy(tmp, c->v0, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      for (int64_t i = 0; i < n; ++i) perm_out[i] = tmp[i];
    delete[] tmp;
    if (e != hipSuccess) return hip_fail(c, e, "copy perm");
  }
  double* dst[3] = {xs, ys, zs};
  const double* dsrc[3] = {c->xs, c->ys, c->zs};
  for (int k = 0; k < 3; ++k)
    if (dst[k]) HIPCHECK(c, hipMemcpy(dst[k], dsrc[k], sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int pepc_gpu_traverse_sinks_device(void* p, int64_t m, const double* x, const double* y,
                                   const double* z, int32_t mode, double* ex, double* ey, double* ez,
                                   double* pot, double* work) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  return traverse_sinks_impl(c, m, x, y, z, mode, ex, ey, ez, pot, work);
}

int pepc_gpu_traverse_sinks_noalloc(void* p, int64_t m, const double* x, const double* y,
                                    const double* z, int32_t mode, double* ex, double* ey,
                                    double* ez, double* pot, double* work) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  int r = traverse_sinks_checks(c, m, x, y, z, mode, ex, ey, ez, pot, work);
  if (r) return r;
  if (m > kNoAllocSinks)
    return fail(c, PEPC_GPU_EINVAL,
                "traverse_sinks_noalloc: " + std::to_string(m) + " sinks exceed the " +
                    std::to_string(kNoAllocSinks) + "-sink scratch (use pepc_gpu_traverse_sinks_device)");
  const int dev = c->cfg.device;
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev >= 0 && dev < kMaxDevices) base = g_scratch_ptr[dev];
    if (!base)
      return fail(c, PEPC_GPU_EHIP,
                  std::string("traverse_sinks_noalloc: no scratch on this device (hipGetSymbolAddress: ") +
                      hipGetErrorString(dev >= 0 && dev < kMaxDevices ? g_scratch_err[dev]
                                                                       : hipErrorInvalidDevice) + ")");
  }
  HIPCHECK(c, hipSetDevice(dev));
  if (m == 0) return 0;
  return traverse_sinks_run(c, scratch_sink_bufs(base), m, x, y, z, mode, ex, ey, ez, pot, work);
}

int pepc_gpu_traverse_sinks(void* p, int64_t m, const double* x, const double* y, const double* z,
                            int32_t mode, double* ex, double* ey, double* ez, double* pot,
                            double* work) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  if (m < 0) return fail(c, PEPC_GPU_EINVAL, "traverse_sinks: bad sink count");
  if (m == 0) return 0;
  if (!x || !y || !z) return fail(c, PEPC_GPU_EINVAL, "traverse_sinks: null argument");
  int r = ensure_sinks(c, m);
  if (r) return r;
  const double* src[3] = {x, y, z};
  for (int k = 0; k < 3; ++k)
    HIPCHECK(c, hipMemcpyAsync(c->sk_io[k], src[k], sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
  r = traverse_sinks_impl(c, m, c->sk_io[0], c->sk_io[1], c->sk_io[2], mode, c->sk_io[3], c->sk_io[4],
                          c->sk_io[5], c->sk_io[6], c->sk_io[7]);
  if (r) return r;
  r = sync_impl(c, nullptr);
  if (r) return r;
  double* dst[5] = {ex, ey, ez, pot, work};
  for (int k = 0; k < 5; ++k)
    if (dst[k]) HIPCHECK(c, hipMemcpy(dst[k], c->sk_io[3 + k], sizeof(double) * m, hipMemcpyDeviceToHost));
  return 0;
}

int pepc_gpu_read_sink_counts(void* p, int64_t cap, uint32_t* p2p, uint32_t* m2p) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  const int64_t m = c->sk_m;
  if (cap < m) return fail(c, PEPC_GPU_EINVAL, "read_sink_counts: capacity below the sink count");
  if (m == 0) return 0;
  std::vector<uint32_t> a(m), b(m), perm;
  HIPCHECK(c, hipMemcpy(a.data(), c->sk_last_p2p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(b.data(), c->sk_last_m2p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  if (c->sk_mode == PEPC_GPU_SINKS_GROUPED) {
    perm.resize(m);
    HIPCHECK(c, hipMemcpy(perm.data(), c->sk_last_perm, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  }
  for (int64_t i = 0; i < m; ++i) {
    const int64_t d = perm.empty() ? i : (int64_t)perm[i];
    if (p2p) p2p[d] = a[i];
    if (m2p) m2p[d] = b[i];
  }
  return 0;
}

int pepc_gpu_check_tree(void* p, int64_t* nbad) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!c->grown) return fail(c, PEPC_GPU_EINVAL, "check_tree: no tree (call grow first)");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = sync_impl(c, nullptr);
  if (r) return r;
  unsigned long long* res = nullptr;
  HIPCHECK(c, hipMalloc((void**)&res, 8 * sizeof(unsigned long long)));
  unsigned long long h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipError_t e = hipMemcpyAsync(res, h, sizeof(h), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch_tree_check(c->stream, c->link, c->nparent, c->n, c->sc, res);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(res);
  if (e != hipSuccess) return hip_fail(c, e, "check_tree");
  const unsigned long long n = (unsigned long long)c->n;
  const DevScalars& st = *c->sc_host;
  int64_t bad = (int64_t)h[3];
  if (h[0] != n) ++bad;                                  // leaves = particles
  if (h[1] != st.ntwig) ++bad;                           // twigs = recorded twigs
  if (h[2] != (unsigned long long)c->nsys && !(c->nsys > 1 && h[2] < (unsigned long long)c->nsys))
    ++bad;                                               // one root per non-empty system
  if (h[4] != n) ++bad;                                  // roots cover every particle
  if (h[5] != n * (n - 1) / 2) ++bad;                    // leaves hold 0..n-1 once
  const unsigned __int128 n128 = n;
  if (h[6] != (unsigned long long)((n128 - 1) * n128 * (2 * n128 - 1) / 6)) ++bad;   // mod 2^64
  if (nbad) *nbad = bad;
  return 0;
}

int pepc_gpu_traverse(void* p, double* ex, double* ey, double* ez, double* pot, double* work) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  if (!c->grown) return fail(c, PEPC_GPU_EINVAL, "traverse: no tree (call grow first)");
  int r = ensure_host_io(c, c->n);
  if (r) return r;
  r = traverse_device_impl(c, c->res[0], c->res[1], c->res[2], c->res[3], c->res[4]);
  if (r) return r;
  r = sync_impl(c, nullptr);
  if (r) return r;
  double* dst[5] = {ex, ey, ez, pot, work};
  for (int k = 0; k < 5; ++k)
    if (dst[k]) HIPCHECK(c, hipMemcpy(dst[k], c->res[k], sizeof(double) * c->n, hipMemcpyDeviceToHost));
  return 0;
}

int pepc_gpu_timber(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  c->grown = false;
  c->walked = false;
  return 0;
}

int pepc_gpu_sorted_device(void* p, const uint64_t** keys, const uint32_t** perm, const double** x,
                           const double** y, const double** z, const double** q) {
  Ctx* c = as_ctx(p);
  if (!c || !c->grown) return fail(c, PEPC_GPU_EINVAL, "sorted_device: no tree");
  if (keys) *keys = reinterpret_cast<const uint64_t*>(c->k0);
  if (perm) *perm = c->v0;
  if (x) *x = c->xs;
  if (y) *y = c->ys;
  if (z) *z = c->zs;
  if (q) *q = c->qs;
  return 0;
}

int pepc_gpu_restore_device(void* p, int64_t count, const double* const* in, double* const* out) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!c->grown) return fail(c, PEPC_GPU_EINVAL, "restore: no tree (call grow first)");
  if (count < 0 || (count > 0 && (!in || !out))) return fail(c, PEPC_GPU_EINVAL, "restore: bad arrays");
  for (int64_t k = 0; k < count; ++k) {
    if (!in[k] || !out[k] || in[k] == out[k])
      return fail(c, PEPC_GPU_EINVAL, "restore: null or in-place array");
  }
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  for (int64_t k = 0; k < count; ++k) launch_scatter_back(c->stream, c->n, c->v0, in[k], out[k]);
  HIPCHECK(c, hipGetLastError());
  return 0;
}

int pepc_gpu_read_sorted(void* p, uint64_t* keys, int64_t* perm, double* x, double* y, double* z,
                         double* q) {
  Ctx* c = as_ctx(p);
  if (!c || !c->grown) return fail(c, PEPC_GPU_EINVAL, "read_sorted: no tree");
  int r = sync_impl(c, nullptr);
  if (r) return r;
  const int64_t n = c->n;
  if (keys) HIPCHECK(c, hipMemcpy(keys, c->k0, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
  if (perm) {
    uint32_t* tmp = new uint32_t[n];
    hipError_t e = hipMemcpy(tmp, c->v0, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      for (int64_t i = 0; i < n; ++i) perm[i] = tmp[i];
    delete[] tmp;
    if (e != hipSuccess) return hip_fail(c, e, "copy perm");
  }
  double* dst[4] = {x, y, z, q};
  const double* src[4] = {c->xs, c->ys, c->zs, c->qs};
  for (int k = 0; k < 4; ++k)
    if (dst[k]) HIPCHECK(c, hipMemcpy(dst[k], src[k], sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int pepc_gpu_read_tree(void* p, int64_t cap, uint32_t* link, uint32_t* first, double* mom,
                       int64_t* nnodes) {
  Ctx* c = as_ctx(p);
  if (!c || !c->grown) return fail(c, PEPC_GPU_EINVAL, "read_tree: no tree");
  int r = sync_impl(c, nullptr);
  if (r) return r;
  const int64_t m = c->sc_host->nnodes;
  if (nnodes) *nnodes = m;
  if (m > cap) return fail(c, PEPC_GPU_EINVAL, "read_tree: capacity too small");
  if (mom) {   // leaves keep no record during a solve; materialise them for the copy
    launch_leaf_moments(c->stream, c->tb);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  if (link) HIPCHECK(c, hipMemcpy(link, c->link, sizeof(NodeLink) * m, hipMemcpyDeviceToHost));
  if (first) HIPCHECK(c, hipMemcpy(first, c->nfirst, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  if (mom) HIPCHECK(c, hipMemcpy(mom, c->mom64, sizeof(double) * 16 * m, hipMemcpyDeviceToHost));
  return 0;
}

int pepc_gpu_read_box(void* p, int64_t sys, double* boxmin, double* boxmax, double* boxsize) {
  Ctx* c = as_ctx(p);
  if (!c || !c->grown) return fail(c, PEPC_GPU_EINVAL, "read_box: no tree");
  if (sys < 0 || sys >= c->nsys) return fail(c, PEPC_GPU_EINVAL, "read_box: no such system");
  int r = sync_impl(c, nullptr);
  if (r) return r;
  DevBox b;
  HIPCHECK(c, hipMemcpy(&b, c->box + sys, sizeof(DevBox), hipMemcpyDeviceToHost));
  for (int d = 0; d < 3; ++d) {
    if (boxmin) boxmin[d] = b.boxmin[d];
    if (boxmax) boxmax[d] = b.boxmax[d];
    if (boxsize) boxsize[d] = b.boxsize[d];
  }
  return 0;
}

int pepc_gpu_read_walk(void* p, int64_t gcap, uint32_t* gfirst, uint32_t* gcount, uint32_t* gmac,
                       uint32_t* p2p, uint32_t* m2p, int64_t* ngroups) {
  Ctx* c = as_ctx(p);
  if (!c || !c->grown) return fail(c, PEPC_GPU_EINVAL, "read_walk: no tree");
  int r = sync_impl(c, nullptr);
  if (r) return r;
  const int64_t g = c->sc_host->ngroups;
  if (ngroups) *ngroups = g;
  if (g > gcap) return fail(c, PEPC_GPU_EINVAL, "read_walk: capacity too small");
  if (gfirst) HIPCHECK(c, hipMemcpy(gfirst, c->gfirst, sizeof(uint32_t) * g, hipMemcpyDeviceToHost));
  if (gcount) HIPCHECK(c, hipMemcpy(gcount, c->gcount, sizeof(uint32_t) * g, hipMemcpyDeviceToHost));
  if (c->walked) {
    if (gmac && g > 0) {
      uint32_t* tmp = new uint32_t[8 * g];
      hipError_t e = hipMemcpy(tmp, c->gstat, sizeof(uint32_t) * 8 * g, hipMemcpyDeviceToHost);
      if (e == hipSuccess)
        for (int64_t i = 0; i < g; ++i) gmac[i] = tmp[8 * i];
      delete[] tmp;
      if (e != hipSuccess) return hip_fail(c, e, "copy group records");
    }
    if (p2p) HIPCHECK(c, hipMemcpy(p2p, c->np2p, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost));
    if (m2p) HIPCHECK(c, hipMemcpy(m2p, c->nm2p, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost));
  }
  return 0;
}

// directforce (src/utils/module_directsum.f90:34-172): with periodic
// boundaries every source in every mirror box, self only in the central box
// (:80-93), then the lattice far field at the targets from all sources
// (calc_force_after_grow + calc_force_per_particle, :130-137), which leaves
// the context's lattice moments at these sources.
static void enqueue_direct(Ctx* c, hipStream_t s, int64_t n, const double* x, const double* y,
                           const double* z, const double* q, int64_t ntest, const int64_t* testidx,
                           double eps2, double sign, double* ex, double* ey, double* ez,
                           double* pot) {
  if (!c->periodic) {
    launch_direct(s, n, x, y, z, q, ntest, testidx, eps2, sign, ex, ey, ez, pot);
    return;
  }
  launch_direct(s, n, x, y, z, q, ntest, testidx, eps2, sign, ex, ey, ez, pot, c->nbox, c->vbox);
  if (c->far_field && ntest > 0) {
    launch_lattice_sources(s, n, x, y, z, q, c->lgeom, c->ml, c->fact, c->lpart, c->ldpart, c->lom,
                           c->ldip, c->mu, c->fict, c->lcoef);
    launch_lattice_field(s, ntest, x, y, z, c->lgeom, c->lcoef, c->linv, c->fict, sign, ex, ey, ez,
                         pot, testidx);
  }
}

int pepc_gpu_direct_device(void* p, int64_t n, const double* x, const double* y, const double* z,
                           const double* q, int64_t ntest, const int64_t* testidx, double eps2,
                           double sign, double* ex, double* ey, double* ez, double* pot) {
  Ctx* c = as_ctx(p);
  if (!c || !x || !y || !z || !q || !testidx || !ex || !ey || !ez || !pot || n < 1 || ntest < 0)
    return fail(c, PEPC_GPU_EINVAL, "direct: bad argument");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  enqueue_direct(c, c->stream, n, x, y, z, q, ntest, testidx, eps2, sign, ex, ey, ez, pot);
  HIPCHECK(c, hipGetLastError());
  return 0;
}

int pepc_gpu_direct(void* p, int64_t n, const double* x, const double* y, const double* z,
                    const double* q, int64_t ntest, const int64_t* testidx, double eps2,
                    double sign, double* ex, double* ey, double* ez, double* pot) {
  Ctx* c = as_ctx(p);
  if (!c || !x || !y || !z || !q || !testidx || !ex || !ey || !ez || !pot || n < 1 || ntest < 0)
    return fail(c, PEPC_GPU_EINVAL, "direct: bad argument");
  for (int64_t j = 0; j < ntest; ++j)
    if (testidx[j] < 0 || testidx[j] >= n) return fail(c, PEPC_GPU_EINVAL, "direct: test index out of range");
  if (ntest == 0) return 0;
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  // a diagnostic, not a solve: one temporary block for sources, indices and results
  const size_t nb = sizeof(double) * (size_t)(4 * n + 5 * ntest);
  double* d = nullptr;
  if (hipMalloc((void**)&d, nb) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, PEPC_GPU_ENOMEM, "direct: no device memory for the sources");
  }
  double* dx = d; double* dy = dx + n; double* dz = dy + n; double* dq = dz + n;
  int64_t* dt = reinterpret_cast<int64_t*>(dq + n);
  double* r = reinterpret_cast<double*>(dt + ntest);
  hipStream_t s = c->stream;
  hipError_t e = hipSuccess;
  const double* src[4] = {x, y, z, q};
  double* dst[4] = {dx, dy, dz, dq};
  for (int k = 0; k < 4 && e == hipSuccess; ++k)
    e = hipMemcpyAsync(dst[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dt, testidx, sizeof(int64_t) * ntest, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    enqueue_direct(c, s, n, dx, dy, dz, dq, ntest, dt, eps2, sign, r, r + ntest, r + 2 * ntest,
                   r + 3 * ntest);
    e = hipGetLastError();
  }
  double* out[4] = {ex, ey, ez, pot};
  for (int k = 0; k < 4 && e == hipSuccess; ++k)
    e = hipMemcpyAsync(out[k], r + k * ntest, sizeof(double) * ntest, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, PEPC_GPU_EHIP, std::string("direct: ") + hipGetErrorString(e));
  return 0;
}

int pepc_gpu_timings(void* p, double* t_grow_ms, double* t_walk_ms) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  float a = 0.f, b = 0.f;
  // an interval whose events were not recorded reads -1; the failed query's
  // error is cleared so that the next call's launch check does not see it
  if (t_grow_ms) {
    *t_grow_ms = -1.0;
    if (c->grown) {
      if (hipEventElapsedTime(&a, c->ev[0], c->ev[1]) == hipSuccess) *t_grow_ms = a;
      else (void)hipGetLastError();
    }
  }
  if (t_walk_ms) {
    *t_walk_ms = -1.0;
    if (c->walked) {
      if (hipEventElapsedTime(&b, c->ev[2], c->ev[3]) == hipSuccess) *t_walk_ms = b;
      else (void)hipGetLastError();
    }
  }
  return 0;
}

// ------------------------------------------------------------ timestep
}  // extern "C"

namespace {

StepState step_state(Ctx* c) {
  StepState st;
  st.x = c->st[0]; st.y = c->st[1]; st.z = c->st[2];
  st.vx = c->st[3]; st.vy = c->st[4]; st.vz = c->st[5];
  st.q = c->st[6]; st.m = c->st[7];
  st.ax = c->st[8]; st.ay = c->st[9]; st.az = c->st[10];
  st.label = c->label;
  st.tvx = c->tmp[0]; st.tvy = c->tmp[1]; st.tvz = c->tmp[2]; st.tm = c->tmp[3];
  st.tlabel = c->tlabel;
  return st;
}

// One kick-drift-kick step (dt = 0: the initial solve, which only computes
// accelerations and energies).  Launches only; no host synchronisation.
int step_impl(Ctx* c, double dt) {
  const int64_t n = c->n;
  const StepState st = step_state(c);
  hipStream_t s = c->stream;
  if (dt != 0.0) launch_kick_drift(s, n, dt, st);
  if (dt != 0.0 && c->periodic) launch_constrain_periodic(s, n, st.x, st.y, st.z, c->lgeom);
  int r = grow_device_impl(c, n, st.x, st.y, st.z, st.q, c->nsys,
                           c->nsys > 1 ? c->hoff.data() : nullptr);
  if (r) return r;
  r = traverse_device_impl(c, c->res[0], c->res[1], c->res[2], c->res[3], c->res[4]);
  if (r) return r;
  s = c->stream;
  launch_reorder_kick(s, n, dt, c->v0, c->xs, c->ys, c->zs, c->qs, c->res[0], c->res[1], c->res[2],
                      st);
  launch_energy(s, (int)c->nsys, c->nechunk, c->echunk, c->esyschunk, c->res[3], st, c->epartial,
                c->ehist, c->ehist_cap, c->ecount);
  HIPCHECK(c, hipGetLastError());
  c->nrec += 1;
  return 0;
}

}  // namespace

extern "C" {

int pepc_gpu_state_set(void* p, int64_t n, int64_t nsys, const int64_t* offsets, const double* x,
                       const double* y, const double* z, const double* vx, const double* vy,
                       const double* vz, const double* q, const double* m, int64_t max_records) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  const double* src[8] = {x, y, z, vx, vy, vz, q, m};
  for (const double* a : src)
    if (!a) return fail(c, PEPC_GPU_EINVAL, "state_set: null array");
  if (n < 1 || n > (int64_t)0x7fffffff) return fail(c, PEPC_GPU_EINVAL, "state_set: bad n");
  if (nsys < 1 || nsys > 65536) return fail(c, PEPC_GPU_EINVAL, "state_set: nsys must be in [1, 65536]");
  if (nsys > 1) {
    if (!offsets) return fail(c, PEPC_GPU_EINVAL, "state_set: null offsets");
    if (offsets[0] != 0 || offsets[nsys] != n)
      return fail(c, PEPC_GPU_EINVAL, "state_set: offsets must start at 0 and end at n");
    for (int64_t k = 0; k < nsys; ++k)
      if (offsets[k + 1] < offsets[k])
        return fail(c, PEPC_GPU_EINVAL, "state_set: offsets must not decrease");
  }
  if (max_records < 1) return fail(c, PEPC_GPU_EINVAL, "state_set: max_records must be >= 1");
  for (int64_t k = 0; k < n; ++k)
    if (!(m[k] > 0.0)) return fail(c, PEPC_GPU_EINVAL, "state_set: masses must be positive");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = ensure_particles(c, n);
  if (r) return r;
  r = ensure_host_io(c, n);
  if (r) return r;
  r = ensure_systems(c, nsys);
  if (r) return r;
  if (nsys > 1) {
    r = set_offsets(c, nsys, offsets);
    if (r) return r;
  }
  for (int k = 0; k < 11; ++k) ALLOC(c, c->st[k], c->ncap);
  for (int k = 0; k < 4; ++k) ALLOC(c, c->tmp[k], c->ncap);
  ALLOC(c, c->label, c->ncap);
  ALLOC(c, c->tlabel, c->ncap);
  ALLOC(c, c->ehist, max_records * nsys * 2);
  ALLOC(c, c->ecount, 64);
  {
    // energy reduction chunks: whole systems cut into pieces of kEnergyChunk
    std::vector<int64_t> cb;
    std::vector<int> sc(nsys + 1);
    for (int64_t k = 0; k < nsys; ++k) {
      const int64_t b = nsys > 1 ? offsets[k] : 0, e = nsys > 1 ? offsets[k + 1] : n;
      sc[k] = (int)cb.size();
      for (int64_t i = b; i < e; i += kEnergyChunk) cb.push_back(i);
    }
    sc[nsys] = (int)cb.size();
    const int nch = (int)cb.size();
    cb.push_back(n);
    ALLOC(c, c->echunk, nch + 1);
    ALLOC(c, c->esyschunk, nsys + 1);
    ALLOC(c, c->epartial, 2 * (int64_t)(nch > 0 ? nch : 1));
    HIPCHECK(c, hipMemcpyAsync(c->echunk, cb.data(), sizeof(int64_t) * (nch + 1), hipMemcpyHostToDevice,
                               c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->esyschunk, sc.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->nechunk = nch;
  }
  c->ehist_cap = max_records;
  drop_graph(c);
  for (int k = 0; k < 8; ++k)
    HIPCHECK(c, hipMemcpyAsync(c->st[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  {
    std::vector<long long> lab(n);
    for (int64_t k = 0; k < n; ++k) lab[k] = k;
    HIPCHECK(c, hipMemcpyAsync(c->label, lab.data(), sizeof(long long) * n, hipMemcpyHostToDevice,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  // accelerations start at zero; the initial solve (dt = 0) sets them
  launch_counter_set(c->stream, c->ecount, 0u);
  c->n = n;
  c->nsys = nsys;
  c->nrec = 0;
  c->has_state = true;
  r = step_impl(c, 0.0);   // initial forces and energies: record 0
  if (r) return r;
  return sync_impl(c, nullptr);
}

int pepc_gpu_run(void* p, int64_t nsteps, double dt, int flags) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!c->has_state) return fail(c, PEPC_GPU_EINVAL, "run: no particle state (call state_set)");
  if (nsteps < 0 || !(dt == dt)) return fail(c, PEPC_GPU_EINVAL, "run: bad nsteps or dt");
  if (flags & ~PEPC_GPU_RUN_GRAPH) return fail(c, PEPC_GPU_EINVAL, "run: unknown flags");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  if (!(flags & PEPC_GPU_RUN_GRAPH)) {
    for (int64_t k = 0; k < nsteps; ++k) {
      int r = step_impl(c, dt);
      if (r) return r;
    }
    return 0;
  }
  if (nsteps == 0) return 0;
  if (!c->gexec || c->graph_dt != dt) {
    drop_graph(c);
    if (!c->cstream) HIPCHECK(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
    // order the capture stream after the work already queued on the context stream
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    hipStream_t user = c->stream;
    c->stream = c->cstream;
    c->capturing = true;
    hipError_t e = hipStreamBeginCapture(c->cstream, hipStreamCaptureModeThreadLocal);
    int r = e == hipSuccess ? step_impl(c, dt) : 0;
    hipGraph_t g = nullptr;
    hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(c->cstream, &g) : e;
    c->capturing = false;
    c->stream = user;
    c->nrec -= 1;   // the captured step did not run
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamBeginCapture");
    if (r) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    if (e2 != hipSuccess) return hip_fail(c, e2, "hipStreamEndCapture");
    c->graph = g;
    HIPCHECK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    c->graph_dt = dt;
  }
  for (int64_t k = 0; k < nsteps; ++k) {
    HIPCHECK(c, hipGraphLaunch(c->gexec, c->stream));
    c->nrec += 1;
  }
  c->walked = true;
  return 0;
}

int pepc_gpu_state_get(void* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                       double* q, double* m, int64_t* label) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!c->has_state) return fail(c, PEPC_GPU_EINVAL, "state_get: no particle state");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = sync_impl(c, nullptr);
  if (r) return r;
  double* dst[8] = {x, y, z, vx, vy, vz, q, m};
  for (int k = 0; k < 8; ++k)
    if (dst[k]) HIPCHECK(c, hipMemcpy(dst[k], c->st[k], sizeof(double) * c->n, hipMemcpyDeviceToHost));
  if (label)
    HIPCHECK(c, hipMemcpy(label, c->label, sizeof(int64_t) * c->n, hipMemcpyDeviceToHost));
  return 0;
}

int pepc_gpu_state_device(void* p, double** arrays, const int64_t** label) {
  Ctx* c = as_ctx(p);
  if (!c || !arrays) return PEPC_GPU_EINVAL;
  if (!c->has_state) return fail(c, PEPC_GPU_EINVAL, "state_device: no particle state");
  for (int k = 0; k < 11; ++k) arrays[k] = c->st[k];
  if (label) *label = reinterpret_cast<const int64_t*>(c->label);
  return 0;
}

}  // extern "C"

namespace {

int ensure_adjoint(Ctx* c) {
  if (c->adj_mcap != c->mcap || c->adj_ncap != c->ncap) {
    ALLOC(c, c->nadj, c->mcap * 16);
    ALLOC(c, c->parent, c->mcap);
    ALLOC(c, c->leafof, c->ncap);
    ALLOC(c, c->aqueue, 64);
    c->adj_mcap = c->mcap;
    c->adj_ncap = c->ncap;
  }
  if (c->list_ncap != c->ncap) {
    ALLOC(c, c->lcount, c->ncap);
    ALLOC(c, c->loff, c->ncap + 1);
    ALLOC(c, c->ltotal, 8);
    c->list_ncap = c->ncap;
  }
  return 0;
}

}  // namespace

extern "C" {

int pepc_gpu_freeze(void* p, int64_t* nentries) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!c->grown) return fail(c, PEPC_GPU_EINVAL, "freeze: no tree (call grow first)");
  if (c->periodic) return fail(c, PEPC_GPU_EINVAL, "freeze: not available with periodic boundaries");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = ensure_adjoint(c);
  if (r) return r;
  r = sync_impl(c, nullptr);
  if (r) return r;
  const int64_t ng = c->sc_host->ngroups;
  ListArgs a;
  a.x = c->xs; a.y = c->ys; a.z = c->zs;
  a.geom = c->geom; a.link = c->link;
  a.gfirst = c->gfirst; a.gcount = c->gcount;
  a.sys = c->nsys > 1 ? c->sys : nullptr;
  a.sysroot = c->sysroot;
  a.sc = c->sc; a.theta2 = c->cfg.theta2;
  a.count = c->lcount; a.off = c->loff; a.entries = c->lent; a.queue = c->aqueue;
  launch_lists(c->stream, a, false, c->walk_blocks, c->ltotal);
  HIPCHECK(c, hipGetLastError());
  unsigned long long total = 0;
  HIPCHECK(c, hipMemcpyAsync(&total, c->ltotal, sizeof(total), hipMemcpyDeviceToHost, c->stream));
  r = sync_impl(c, nullptr);
  if (r) return r;
  if (total > 0xFFFFFFFFull)
    return fail(c, PEPC_GPU_EOVERFLOW, "freeze: interaction lists exceed 2^32 entries");
  launch_scan(c->stream, c->lcount, c->loff, ng, c->bsum, nullptr);
  if ((int64_t)total > c->lent_cap) {
    ALLOC(c, c->lent, (int64_t)total);
    c->lent_cap = (int64_t)total;
  }
  a.entries = c->lent;
  launch_lists(c->stream, a, true, c->walk_blocks, c->ltotal);
  HIPCHECK(c, hipGetLastError());
  r = sync_impl(c, nullptr);
  if (r) return r;
  c->frozen = true;
  c->frozen_entries = (int64_t)total;
  if (nentries) *nentries = (int64_t)total;
  return 0;
}

int pepc_gpu_replay_device(void* p, const double* x, const double* y, const double* z,
                           const double* q, double* ex, double* ey, double* ez, double* pot) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!x || !y || !z || !q || !ex || !ey || !ez || !pot)
    return fail(c, PEPC_GPU_EINVAL, "replay: null argument");
  if (!c->frozen) return fail(c, PEPC_GPU_EINVAL, "replay: no frozen lists (call freeze after grow)");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  hipStream_t s = c->stream;
  launch_gather4(s, c->n, c->v0, x, y, z, q, c->xs, c->ys, c->zs, c->qs);
  launch_leaf_geom(s, c->link, c->xs, c->ys, c->zs, c->qs, c->geom, c->mcap, c->sc);
  launch_upward(s, c->tb);
  ReplayArgs a;
  a.x = c->xs; a.y = c->ys; a.z = c->zs;
  a.geom = c->geom; a.link = c->link; a.mom64 = c->mom64;
  a.gfirst = c->gfirst; a.gcount = c->gcount;
  a.count = c->lcount; a.off = c->loff; a.entries = c->lent;
  a.sc = c->sc; a.eps2 = c->cfg.eps2; a.sign = c->cfg.sign;
  a.ex = ex; a.ey = ey; a.ez = ez; a.pot = pot;
  a.queue = c->aqueue;
  launch_replay(s, a, c->walk_blocks);
  HIPCHECK(c, hipGetLastError());
  return 0;
}

int pepc_gpu_adjoint_device(void* p, const double* gex, const double* gey, const double* gez,
                            const double* gpot, double* gx, double* gy, double* gz, double* gq) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!gex || !gey || !gez || !gpot || !gx || !gy || !gz || !gq)
    return fail(c, PEPC_GPU_EINVAL, "adjoint: null argument");
  if (!c->grown) return fail(c, PEPC_GPU_EINVAL, "adjoint: no tree (call grow first)");
  if (c->periodic) return fail(c, PEPC_GPU_EINVAL, "adjoint: not available with periodic boundaries");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = ensure_adjoint(c);
  if (r) return r;
  AdjointArgs a;
  a.n = c->n; a.cap = c->mcap;
  a.x = c->xs; a.y = c->ys; a.z = c->zs; a.q = c->qs;
  a.geom = c->geom; a.link = c->link; a.mom64 = c->mom64;
  a.gfirst = c->gfirst; a.gcount = c->gcount;
  a.sys = c->nsys > 1 ? c->sys : nullptr;
  a.sysroot = c->sysroot;
  a.sc = c->sc;
  a.theta2 = c->cfg.theta2; a.eps2 = c->cfg.eps2; a.sign = c->cfg.sign;
  a.gex = gex; a.gey = gey; a.gez = gez; a.gpot = gpot;
  a.gx = gx; a.gy = gy; a.gz = gz; a.gq = gq;
  a.nadj = c->nadj; a.parent = c->parent; a.leafof = c->leafof; a.queue = c->aqueue;
  launch_adjoint(c->stream, a, c->walk_blocks);
  HIPCHECK(c, hipGetLastError());
  return 0;
}

int pepc_gpu_get_config(const void* p, pepc_gpu_config* cfg) {
  const Ctx* c = reinterpret_cast<const Ctx*>(p);
  if (!c || !cfg) return PEPC_GPU_EINVAL;
  *cfg = c->cfg;
  cfg->walk_blocks = c->walk_blocks;
  return 0;
}

int pepc_gpu_autotune(void* p, int64_t n, int64_t nsys, const int64_t* offsets, const double* x,
                      const double* y, const double* z, const double* q, int32_t reps,
                      pepc_gpu_tune_entry* out, int64_t cap, int64_t* ncand) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!x || !y || !z || !q || n < 1 || reps < 1 || nsys < 1 || (nsys > 1 && !offsets))
    return fail(c, PEPC_GPU_EINVAL, "autotune: bad argument");
  if (nsys > 1 && offsets[nsys] != n) return fail(c, PEPC_GPU_EINVAL, "autotune: offsets must end at n");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = ensure_particles(c, n);
  if (r) return r;
  r = ensure_host_io(c, n);
  if (r) return r;
  const int prec = c->cfg.precision == PEPC_GPU_FP64 ? kWalkFP64 : kWalkFP32Kahan;
  const int tiles[2] = {prec == kWalkFP64 ? 32 : 64, prec == kWalkFP64 ? 64 : 128};
  const int sizes[3] = {16, 32, 64};
  const int packs[3] = {0, 32, 64};
  struct Cand { int g, pk, lt, wb; double tg, tw; };
  std::vector<Cand> cands;
  for (int g : sizes)
    for (int pk : packs)
      for (int lt : tiles) {
        const int full = walk_default_blocks(prec, lt);
        for (int wb : {full, full / 2}) cands.push_back({g, pk, lt, wb > 0 ? wb : 1, 0.0, 0.0});
      }
  const pepc_gpu_config saved = c->cfg;
  const int saved_wb = c->walk_blocks;
  const int64_t ns = nsys > 1 ? nsys : 1;
  int best = -1;
  for (size_t k = 0; k < cands.size(); ++k) {
    Cand& cd = cands[k];
    c->cfg.group_size = cd.g;
    c->cfg.group_pack = cd.pk;
    c->cfg.list_tile = cd.lt;
    c->walk_blocks = cd.wb;
    double tbest = 1e300;
    for (int rep = 0; rep <= reps; ++rep) {   // rep 0 warms up
      r = grow_device_impl(c, n, x, y, z, q, ns, nsys > 1 ? offsets : nullptr);
      if (!r) r = traverse_device_impl(c, c->res[0], c->res[1], c->res[2], c->res[3], c->res[4]);
      if (!r) r = sync_impl(c, nullptr);
      float tg = 0.f, tw = 0.f;
      if (!r && (hipEventElapsedTime(&tg, c->ev[0], c->ev[1]) != hipSuccess ||
                 hipEventElapsedTime(&tw, c->ev[2], c->ev[3]) != hipSuccess))
        r = fail(c, PEPC_GPU_EHIP, "autotune: event timing failed");
      if (r) {
        c->cfg = saved;
        c->walk_blocks = saved_wb;
        c->grown = false;
        return r;
      }
      if (rep > 0 && (double)tg + (double)tw < tbest) {
        tbest = (double)tg + (double)tw;
        cd.tg = tg;
        cd.tw = tw;
      }
    }
    if (best < 0 || cd.tg + cd.tw < cands[best].tg + cands[best].tw) best = (int)k;
  }
  c->cfg = saved;
  c->cfg.group_size = cands[best].g;
  c->cfg.group_pack = cands[best].pk;
  c->cfg.list_tile = cands[best].lt;
  c->cfg.walk_blocks = cands[best].wb;
  c->walk_blocks = cands[best].wb;
  c->grown = false;   // the next grow builds groups with the chosen configuration
  c->walked = false;
  if (ncand) *ncand = (int64_t)cands.size();
  if (out) {
    // best first, then the candidates in the order tried
    int64_t j = 0;
    if (j < cap) {
      const Cand& b = cands[best];
      out[j++] = {b.g, b.pk, b.lt, b.wb, b.tg, b.tw};
    }
    for (size_t k = 0; k < cands.size() && j < cap; ++k) {
      const Cand& d = cands[k];
      out[j++] = {d.g, d.pk, d.lt, d.wb, d.tg, d.tw};
    }
  }
  return 0;
}

int pepc_gpu_read_energies(void* p, int64_t cap, double* out, int64_t* nrec) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  if (!c->has_state) return fail(c, PEPC_GPU_EINVAL, "read_energies: no particle state");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  int r = sync_impl(c, nullptr);
  if (r) return r;
  const int64_t k = c->nrec < c->ehist_cap ? c->nrec : c->ehist_cap;
  if (nrec) *nrec = c->nrec;
  if (out && k > 0) {
    const int64_t m = (k < cap ? k : cap) * c->nsys * 2;
    HIPCHECK(c, hipMemcpy(out, c->ehist, sizeof(double) * m, hipMemcpyDeviceToHost));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Periodic boundaries
}  // extern "C"

namespace {

// factorial of module_fmm_framework.f90:1247-1319: the exact integers up to
// 24! (rounded to double), gamma(n + 1) beyond.
double ref_factorial(int n) {
  if (n <= 24) {
    unsigned __int128 f = 1;
    for (int k = 2; k <= n; ++k) f *= (unsigned)k;
    return (double)f;
  }
  return std::tgamma((double)n + 1.0);
}

bool invert3(const double a[9], double out[9]) {
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8],
               c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
  const double id = 1.0 / det;
  out[0] = c00 * id; out[1] = (a[2] * a[7] - a[1] * a[8]) * id; out[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  out[3] = c01 * id; out[4] = (a[0] * a[8] - a[2] * a[6]) * id; out[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  out[6] = c02 * id; out[7] = (a[1] * a[6] - a[0] * a[7]) * id; out[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  return true;
}

// Lattice indices of the neighbour boxes in the order of
// neighbour_boxes_prepare (module_mirror_boxes.f90:119-149): a box and its
// central-symmetric partner together, the central box last.
std::vector<int> neighbour_boxes(const int per[3], int layers) {
  int sw[3];
  for (int d = 0; d < 3; ++d) sw[d] = per[d] ? layers : 0;
  std::vector<int> b;
  auto put = [&](int i, int j, int k) { b.push_back(i); b.push_back(j); b.push_back(k); };
  for (int i = -sw[0]; i <= -1; ++i)
    for (int j = -sw[1]; j <= sw[1]; ++j)
      for (int k = -sw[2]; k <= sw[2]; ++k) { put(i, j, k); put(-i, -j, -k); }
  for (int j = -sw[1]; j <= -1; ++j)
    for (int k = -sw[2]; k <= sw[2]; ++k) { put(0, j, k); put(0, -j, -k); }
  for (int k = -sw[2]; k <= -1; ++k) { put(0, 0, k); put(0, 0, -k); }
  put(0, 0, 0);
  return b;
}

}  // namespace

extern "C" {

void pepc_gpu_periodic_default(pepc_gpu_periodic* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->mirror_box_layers = 1;
  p->t_lattice[0] = p->t_lattice[4] = p->t_lattice[8] = 1.0;
  p->far_field = 1;
  p->extrinsic_correction = PEPC_GPU_EXTRINSIC_FICTCHARGE;
}

int pepc_gpu_set_periodic(void* ctx, const pepc_gpu_periodic* p) {
  Ctx* c = as_ctx(ctx);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  if (c->capturing) return fail(c, PEPC_GPU_EINVAL, "set_periodic: not during graph capture");
  drop_graph(c);
  if (!p || !(p->periodicity[0] || p->periodicity[1] || p->periodicity[2])) {
    c->periodic = false;
    c->far_field = false;
    c->nbox = 1;
    return 0;
  }
  const int layers = p->mirror_box_layers;
  if (layers < 1 || layers > 3)
    return fail(c, PEPC_GPU_EINVAL, "set_periodic: mirror_box_layers must be in [1, 3]");
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(p->t_lattice[k])) return fail(c, PEPC_GPU_EINVAL, "set_periodic: non-finite lattice");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(p->origin[k])) return fail(c, PEPC_GPU_EINVAL, "set_periodic: non-finite origin");
  LatticeGeom g;
  memcpy(g.lat, p->t_lattice, sizeof(g.lat));
  if (!invert3(g.lat, g.latinv)) return fail(c, PEPC_GPU_EINVAL, "set_periodic: degenerate lattice");
  // The far field runs for any periodicity, as the reference's does
  // (do_periodic .and. include_far_field_if_periodic, coulomb/
  // module_interaction_specific.f90:199-225): the lattice sums, the mirror
  // boxes and the fictitious charges cover the periodic axes only.
  const bool all3 = p->periodicity[0] && p->periodicity[1] && p->periodicity[2];
  const bool far = p->far_field != 0;
  const int corr = p->extrinsic_correction;
  if (far && corr != PEPC_GPU_EXTRINSIC_NONE && corr != PEPC_GPU_EXTRINSIC_REDLACK &&
      corr != PEPC_GPU_EXTRINSIC_FICTCHARGE)
    return fail(c, PEPC_GPU_EINVAL,
                "set_periodic: extrinsic_correction must be NONE (0), REDLACK (1) or FICTCHARGE (2)");
  g.correction = corr;
  g.volume = std::fabs((g.lat[1] * g.lat[5] - g.lat[2] * g.lat[4]) * g.lat[6] +
                       (g.lat[2] * g.lat[3] - g.lat[0] * g.lat[5]) * g.lat[7] +
                       (g.lat[0] * g.lat[4] - g.lat[1] * g.lat[3]) * g.lat[8]);
  for (int d = 0; d < 3; ++d) {
    g.origin[d] = p->origin[d];
    g.center[d] = 0.5 * (g.lat[d] + g.lat[3 + d] + g.lat[6 + d]) + g.origin[d];
    g.per[d] = p->periodicity[d] ? 1 : 0;
  }
  const double* t = g.lat;   // system_uses_principal_axes, module_mirror_boxes.f90:224-236
  g.simple = (t[0] > 0.0 && t[1] == 0.0 && t[2] == 0.0 && t[3] == 0.0 && t[4] > 0.0 && t[5] == 0.0 &&
              t[6] == 0.0 && t[7] == 0.0 && t[8] > 0.0) ? 1 : 0;
  const std::vector<int> bx = neighbour_boxes(g.per, layers);
  const int nb = (int)bx.size() / 3;
  std::vector<double> vec(3 * (size_t)nb);
  for (int b = 0; b < nb; ++b)   // lattice_vect, module_mirror_boxes.f90:73-79
    for (int d = 0; d < 3; ++d)
      vec[3 * b + d] = (double)bx[3 * b] * t[d] + (double)bx[3 * b + 1] * t[3 + d] +
                       (double)bx[3 * b + 2] * t[6 + d];
  c->periodic = false;              // until the new lattice is complete
  c->far_field = false;
  hipStream_t s = c->stream;
  std::vector<double> f;            // host sources stay alive until the final sync
  std::vector<double2> mlh;
  ALLOC(c, c->vbox, 3 * (int64_t)nb);
  HIPCHECK(c, hipMemcpyAsync(c->vbox, vec.data(), sizeof(double) * vec.size(), hipMemcpyHostToDevice, s));
  if (far) {
    ALLOC(c, c->fact, kLatticeFact);
    ALLOC(c, c->ml, kLatticeNT); ALLOC(c, c->mstar, kLatticeNM); ALLOC(c, c->lstar, kLatticeNT);
    ALLOC(c, c->mu, kLatticeNM); ALLOC(c, c->lpart, (int64_t)kLatticeMaxBlocks * kLatticeNM);
    ALLOC(c, c->ldpart, (int64_t)kLatticeMaxBlocks * 4); ALLOC(c, c->fict, kLatticeFict);
    ALLOC(c, c->linv, kLatticeNM + kLatticeLM + 1); ALLOC(c, c->ldip, 4);
    ALLOC(c, c->lom, kLatticeNM); ALLOC(c, c->lcoef, 8 * kLatticeNM);
    launch_lattice_inv(s, c->linv);
    f.resize(kLatticeFact);
    for (int k = 0; k < kLatticeFact; ++k) f[k] = ref_factorial(k);
    HIPCHECK(c, hipMemcpyAsync(c->fact, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice, s));
    auto nsq = [](const double* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; };
    // system_is_unit_cube (module_mirror_boxes.f90:248-255): the reference
    // then loads its pretabulated coefficients (module_fmm_framework.f90:114-121).
    // Those are the 3-D periodic lattice beyond one mirror layer, so with more
    // layers or fewer periodic axes the coefficients are computed for the
    // lattice set (the reference loads the table for any layer count and
    // periodicity, and warns that it holds for the 3-D periodic unit box only).
    const bool unit = layers == 1 && all3 && g.simple && nsq(t) == 1.0 && nsq(t + 3) == 1.0 &&
                      nsq(t + 6) == 1.0;
    if (unit) {
      mlh.assign(kLatticeNT, make_double2(0.0, 0.0));
      for (int k = 0; k < kUnitCubeEntries; ++k) {
        const UnitCubeEntry& e = kUnitCubeML[k];
        mlh[e.l * (e.l + 1) / 2 + e.m] = make_double2(e.v, 0.0);
      }
      HIPCHECK(c, hipMemcpyAsync(c->ml, mlh.data(), sizeof(double2) * mlh.size(), hipMemcpyHostToDevice, s));
    } else {
      launch_lattice_coefficients(s, c->vbox, nb, (double)(2 * layers + 1), c->fact, c->mstar,
                                  c->lstar, 32, c->ml);
      HIPCHECK(c, hipGetLastError());
    }
  }
  // Fictitious charges (module_fmm_framework.f90:535-554): charge p of a box
  // sits at origin + t_p, the last at the origin, and every neighbour box
  // carries the set.  Merged by lattice point; where all of them meet they
  // cancel (the last is minus the sum of the others) and are left out.
  std::vector<int> fpt_h;
  g.fpt = nullptr;
  g.fpos = nullptr;
  g.nfpt = 0;
  if (far && corr == PEPC_GPU_EXTRINSIC_FICTCHARGE) {
    std::vector<std::array<int, 3>> off;
    for (int q = 0; q < 3; ++q) {
      if (!g.per[q]) continue;
      std::array<int, 3> e = {0, 0, 0};
      e[q] = 1;
      off.push_back(e);
    }
    off.push_back({0, 0, 0});
    const int nf = (int)off.size(), all = (1 << nf) - 1;
    std::map<std::array<int, 3>, int> mask;   // ordered: a fixed point order
    for (int b = 0; b < nb; ++b)
      for (int f = 0; f < nf; ++f) {
        const std::array<int, 3> pt = {bx[3 * b] + off[f][0], bx[3 * b + 1] + off[f][1],
                                       bx[3 * b + 2] + off[f][2]};
        mask[pt] |= 1 << f;
      }
    for (const auto& kv : mask) {
      if (kv.second == all) continue;
      fpt_h.insert(fpt_h.end(), {kv.first[0], kv.first[1], kv.first[2], kv.second});
    }
    const int np = (int)fpt_h.size() / 4;
    ALLOC(c, c->lfpt, 4 * (int64_t)np);
    ALLOC(c, c->lfpos, 4 * (int64_t)np);
    HIPCHECK(c, hipMemcpyAsync(c->lfpt, fpt_h.data(), sizeof(int) * fpt_h.size(), hipMemcpyHostToDevice, s));
    g.fpt = c->lfpt;
    g.fpos = c->lfpos;
    g.nfpt = np;
  }
  HIPCHECK(c, hipStreamSynchronize(s));
  c->per = *p;
  c->lgeom = g;
  c->nbox = nb;
  c->periodic = true;
  c->far_field = far;
  return 0;
}

int pepc_gpu_lattice_get(void* ctx, double* ml, double* mu, double* fict, int32_t* nfict) {
  Ctx* c = as_ctx(ctx);
  if (!c) return PEPC_GPU_EINVAL;
  if (!c->periodic || !c->far_field) return fail(c, PEPC_GPU_EINVAL, "lattice_get: no far field set");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (ml) HIPCHECK(c, hipMemcpy(ml, c->ml, sizeof(double2) * kLatticeNT, hipMemcpyDeviceToHost));
  if (mu) HIPCHECK(c, hipMemcpy(mu, c->mu, sizeof(double2) * kLatticeNM, hipMemcpyDeviceToHost));
  double f[kLatticeFict];
  if (fict || nfict) HIPCHECK(c, hipMemcpy(f, c->fict, sizeof(f), hipMemcpyDeviceToHost));
  if (fict) memcpy(fict, f, sizeof(double) * 16);
  if (nfict) *nfict = (int32_t)f[16];
  return 0;
}

int pepc_gpu_constrain_periodic_device(void* ctx, int64_t n, double* x, double* y, double* z) {
  Ctx* c = as_ctx(ctx);
  if (!c) return PEPC_GPU_EINVAL;
  if (n < 0 || (n > 0 && (!x || !y || !z))) return fail(c, PEPC_GPU_EINVAL, "constrain_periodic: bad argument");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  if (!c->periodic || n == 0) return 0;
  launch_constrain_periodic(c->stream, n, x, y, z, c->lgeom);
  HIPCHECK(c, hipGetLastError());
  return 0;
}

}  // extern "C"
