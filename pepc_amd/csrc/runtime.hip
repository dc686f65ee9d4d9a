// runtime.hip -- context, device memory and the C-ABI of include/pepc_gpu.h.
//
// One context owns the device buffers of one particle system: particle
// arrays sized for n (grown, never shrunk, so a timestep loop allocates only
// on its first step), node arrays sized node_factor * n, and the small
// device scalar block the kernels report through.  grow and traverse are
// sequences of kernel launches on the context's stream with no host round
// trip; the host reads the scalars once, in pepc_gpu_sync().
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/pepc_gpu.h"
#include "kernels.h"

using namespace pepc;

namespace {

struct Ctx {
  pepc_gpu_config cfg;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  int64_t n = 0, ncap = 0, mcap = 0, io_cap = 0;
  bool grown = false, walked = false;
  int walk_blocks = 0;

  // particles
  double *in[4] = {nullptr, nullptr, nullptr, nullptr};   // host-path input copies
  unsigned long long *k0 = nullptr, *k1 = nullptr, *oldkeys = nullptr;
  uint32_t *v0 = nullptr, *v1 = nullptr;
  double *xs = nullptr, *ys = nullptr, *zs = nullptr, *qs = nullptr;
  uint8_t *lo = nullptr, *hi = nullptr;
  uint32_t *cnt = nullptr, *gmark = nullptr, *dflag = nullptr, *dlist = nullptr;
  long long* kd = nullptr;
  uint32_t *gfirst = nullptr, *gcount = nullptr, *gstat = nullptr, *queue = nullptr;
  uint32_t *np2p = nullptr, *nm2p = nullptr;
  double *res[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // host-path outputs
  // nodes
  unsigned long long *dkey[2] = {nullptr, nullptr};
  uint32_t *dval[2] = {nullptr, nullptr};
  uint32_t *dpart = nullptr, *inv = nullptr, *flag = nullptr, *fc_rank = nullptr,
           *node_rank = nullptr, *nfirst = nullptr;
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
};

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
                  c->res[2], c->res[3], c->res[4], c->dkey[0], c->dkey[1], c->dval[0], c->dval[1],
                  c->dpart, c->inv, c->flag, c->fc_rank, c->node_rank, c->nfirst, c->link, c->geom,
                  c->mom64, c->wm32, c->wm64, c->hist, c->bsum, c->partial, c->box, c->sc};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

int ensure_particles(Ctx* c, int64_t n) {
  if (n <= c->ncap) return 0;
  const int64_t nc = n + n / 8 + 64;
  ALLOC(c, c->k0, nc); ALLOC(c, c->k1, nc); ALLOC(c, c->oldkeys, nc);
  ALLOC(c, c->v0, nc); ALLOC(c, c->v1, nc);
  ALLOC(c, c->xs, nc); ALLOC(c, c->ys, nc); ALLOC(c, c->zs, nc); ALLOC(c, c->qs, nc);
  ALLOC(c, c->lo, nc); ALLOC(c, c->hi, nc);
  ALLOC(c, c->cnt, nc); ALLOC(c, c->gmark, nc); ALLOC(c, c->dflag, nc); ALLOC(c, c->dlist, nc);
  ALLOC(c, c->kd, nc);
  ALLOC(c, c->gfirst, nc); ALLOC(c, c->gcount, nc); ALLOC(c, c->gstat, 8 * nc);
  ALLOC(c, c->np2p, nc); ALLOC(c, c->nm2p, nc);
  const int64_t mc = (int64_t)(c->cfg.node_factor * (double)nc) + 64;
  ALLOC(c, c->dkey[0], mc); ALLOC(c, c->dkey[1], mc);
  ALLOC(c, c->dval[0], mc); ALLOC(c, c->dval[1], mc);
  ALLOC(c, c->dpart, mc); ALLOC(c, c->inv, mc); ALLOC(c, c->flag, mc);
  ALLOC(c, c->fc_rank, mc + 1); ALLOC(c, c->node_rank, mc); ALLOC(c, c->nfirst, mc);
  ALLOC(c, c->link, mc); ALLOC(c, c->geom, mc); ALLOC(c, c->mom64, mc * 16);
  ALLOC(c, c->wm32, mc); ALLOC(c, c->wm64, mc);
  // radix histogram: 256 digits x sort blocks of the largest array sorted
  const int64_t hc = 256 * sort_blocks(mc > nc ? mc : nc);
  ALLOC(c, c->hist, hc);
  const int64_t bc = scan_blocks(hc > mc ? hc : mc) + 64;
  ALLOC(c, c->bsum, bc);
  c->ncap = nc;
  c->mcap = mc;
  return 0;
}

int ensure_host_io(Ctx* c, int64_t n) {
  // device staging buffers of the host-pointer entry points
  if (n <= c->io_cap) return 0;
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

int grow_device_impl(Ctx* c, int64_t n, const double* x, const double* y, const double* z,
                     const double* q) {
  if (n < 1) return fail(c, PEPC_GPU_EINVAL, "grow: need at least one particle");
  if (n > (int64_t)0x7fffffff) return fail(c, PEPC_GPU_EINVAL, "grow: n exceeds 2^31-1");
  int r = ensure_particles(c, n);
  if (r) return r;
  hipStream_t s = c->stream;
  HIPCHECK(c, hipEventRecord(c->ev[0], s));
  reset_scalars_kernel<<<1, 64, 0, s>>>(c->sc);
  launch_box(s, n, x, y, z, c->cfg.force_cubic_domain, c->partial, c->box, c->sc);
  launch_keys(s, n, x, y, z, c->box, c->k0, c->v0);
  launch_sort_keys(s, c->k0, c->v0, c->k1, c->v1, n, c->hist, c->bsum);
  launch_gather4(s, n, c->v0, x, y, z, q, c->xs, c->ys, c->zs, c->qs);
  launch_dup_detect(s, n, c->k0, c->sc);
  launch_dupfix(s, n, c->k0, c->xs, c->ys, c->zs, c->box, c->kd, c->oldkeys, c->dflag, c->dlist,
                c->bsum, c->sc);
  TreeBuffers t;
  t.n = n;
  t.cap = c->mcap;
  t.keys = c->k0;
  t.x = c->xs; t.y = c->ys; t.z = c->zs; t.q = c->qs;
  t.lo = c->lo; t.hi = c->hi; t.cnt = c->cnt;
  t.dkey[0] = c->dkey[0]; t.dkey[1] = c->dkey[1];
  t.dval[0] = c->dval[0]; t.dval[1] = c->dval[1];
  t.dpart = c->dpart; t.inv = c->inv; t.flag = c->flag; t.fc_rank = c->fc_rank;
  t.node_rank = c->node_rank; t.nfirst = c->nfirst; t.link = c->link; t.geom = c->geom;
  t.mom64 = c->mom64; t.wm32 = c->wm32; t.wm64 = c->wm64;
  t.gmark = c->gmark; t.gfirst = c->gfirst; t.gcount = c->gcount;
  t.hist = c->hist; t.bsum = c->bsum; t.box = c->box; t.sc = c->sc;
  t.prec = c->cfg.precision == PEPC_GPU_FP64 ? kWalkFP64 : kWalkFP32Kahan;
  c->tb = t;
  launch_tree_build(s, t, c->cfg.group_size, c->cfg.group_pack);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipEventRecord(c->ev[1], s));
  c->n = n;
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
  a.geom = c->geom; a.link = c->link; a.wm32 = c->wm32; a.wm64 = c->wm64;
  a.gfirst = c->gfirst; a.gcount = c->gcount; a.sc = c->sc;
  a.theta2 = c->cfg.theta2; a.eps2 = c->cfg.eps2; a.sign = c->cfg.sign;
  a.ex = ex; a.ey = ey; a.ez = ez; a.pot = pot; a.work = work;
  a.np2p = c->np2p; a.nm2p = c->nm2p; a.gstat = c->gstat; a.queue = c->queue;
  launch_walk_reset(s, a);
  HIPCHECK(c, hipEventRecord(c->ev[2], s));
  launch_walk(s, a, c->cfg.precision, c->walk_blocks);
  HIPCHECK(c, hipEventRecord(c->ev[3], s));
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
  if (cfg->mac_select != 0 || cfg->group_size < 1 || cfg->group_size > 64 ||
      cfg->group_pack < 0 || cfg->group_pack > 64 ||
      !(cfg->theta2 >= 0.0) || !(cfg->eps2 >= 0.0) || !(cfg->node_factor >= 1.0) ||
      (cfg->precision != PEPC_GPU_FP32_KAHAN && cfg->precision != PEPC_GPU_FP64) ||
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
      dalloc(c, &c->queue, 64) ||
      hipHostMalloc((void**)&c->sc_host, sizeof(DevScalars), hipHostMallocDefault) != hipSuccess) {
    free_all(c);
    delete c;
    return PEPC_GPU_ENOMEM;
  }
  c->walk_blocks = cfg->walk_blocks > 0 ? cfg->walk_blocks : walk_default_blocks();
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
  return grow_device_impl(c, n, x, y, z, q);
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
  r = grow_device_impl(c, n, c->in[0], c->in[1], c->in[2], c->in[3]);
  if (r) return r;
  r = sync_impl(c, nullptr);
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

int pepc_gpu_direct_device(void* p, int64_t n, const double* x, const double* y, const double* z,
                           const double* q, int64_t ntest, const int64_t* testidx, double eps2,
                           double sign, double* ex, double* ey, double* ez, double* pot) {
  Ctx* c = as_ctx(p);
  if (!c || !x || !y || !z || !q || !testidx || !ex || !ey || !ez || !pot || n < 1 || ntest < 0)
    return fail(c, PEPC_GPU_EINVAL, "direct: bad argument");
  HIPCHECK(c, hipSetDevice(c->cfg.device));
  launch_direct(c->stream, n, x, y, z, q, ntest, testidx, eps2, sign, ex, ey, ez, pot);
  HIPCHECK(c, hipGetLastError());
  return 0;
}

int pepc_gpu_timings(void* p, double* t_grow_ms, double* t_walk_ms) {
  Ctx* c = as_ctx(p);
  if (!c) return PEPC_GPU_EINVAL;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  float a = 0.f, b = 0.f;
  if (t_grow_ms) {
    *t_grow_ms = -1.0;
    if (c->grown && hipEventElapsedTime(&a, c->ev[0], c->ev[1]) == hipSuccess) *t_grow_ms = a;
  }
  if (t_walk_ms) {
    *t_walk_ms = -1.0;
    if (c->walked && hipEventElapsedTime(&b, c->ev[2], c->ev[3]) == hipSuccess) *t_walk_ms = b;
  }
  return 0;
}

}  // extern "C"
