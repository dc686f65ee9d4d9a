// walk.hip -- persistent group walk with P2P/M2P force and potential kernels.
//
// Semantics (module_tree_walk_simple.f90:256-384): a group of target
// particles (a walk tile) descends the tree from the root; a leaf interacts
// with every target (P2P, self excluded); a twig is accepted for M2P only if
// every target passes the Barnes-Hut test theta2*d^2 > b^2(level)
// (coulomb/module_interaction_specific.f90:256-275), otherwise it is opened.
// Interaction counts per target and MAC-evaluation counts per group equal the
// reference's at the same group rule.
//
// Mapping to CDNA4.  One 64-lane wavefront owns one group at a time and pulls
// the next group from a device queue (persistent kernel).  A group of cnt
// targets occupies T = pow2 >= cnt lanes; the S = 64/T copies ("slices")
// split every interaction list between them, so small groups still fill the
// wave.  Traversal is breadth-parallel: each step pops up to 64 nodes from an
// LDS stack, one node per lane, and decides it against the group's bounding
// box; only nodes whose decision the box cannot settle run the exact
// per-target FP64 test (all lanes, one ballot).  Accepted nodes and leaves are
// appended to M2P / P2P lists staged in LDS and evaluated in batches: every
// lane reads the same entry (LDS broadcast) and evaluates it for its target.
//
// Precision.  FP32 path: coordinates relative to the group's first particle
// (differences taken in FP64 once per node, then rounded), FP32 kernels with
// fused multiply-add, per-target compensated (Kahan) accumulation of each
// batch's partial sums.  FP64 path: absolute coordinates, FP64 kernels,
// plain FP64 accumulation, as the reference.
//
// Periodic systems (PER): the group walks the tree once per lattice vector
// vbox of the mirror boxes, with its targets shifted to x - vbox
// (module_libpepc_main.f90:184-187, module_tree_walk_simple.f90:328,345);
// sums run on across the boxes.  Self is excluded in the central box only.
// FP64 keeps separations as (x - vbox) - coc and flushes its lists at the end
// of each box; FP32 puts the shift into the node frame instead (node entries
// relative to first particle - vbox), so its lists carry over.
//
// Sinks (SINK): targets other than the grown particles (the reference's
// traversal of another particle array, module_pepc.f90:394-433).  Targets
// come from their own arrays, one per group (the reference's rule for
// particles without keys, module_tree_walk_simple.f90:155-159) or in given
// runs; self is the reference's test, a zero separation from the leaf's
// particle (:331-340), checked exactly in FP64 against the source
// coordinates; lists are flushed at the end of each box.
//
// The M2P evaluation regroups the reference's expression
// (module_coulomb_kernels.f90:60-145) around the quadratic form
// S = d.A.d of the quadrupole tensor; it is the same polynomial in d and the
// moments (phi = q/r - d.p/r^3 + 3 S/r^5 - tr(A)/r^3 and E = -grad phi with
// r^2 = d^2 + eps^2), evaluated with fewer operations.
#include "kernels.h"

namespace pepc {

namespace {

constexpr int kStackCap = 512;      // entries the pop policy keeps below (FP64)
constexpr int kStackCap32 = 768;    // the same for FP32, whose lists leave LDS for it
constexpr int kStackSlack = 192;    // >= 7 * 21 + 8: depth-first overrun bound
constexpr int kListP = 64;
// Stack entries carry the node's kind in their top bit (set for a twig, and
// for a root whatever it is), so a leaf's step skips the walk-moment loads.
constexpr uint32_t kTwigBit = 0x80000000u;

template <typename R> struct Prec;
template <> struct Prec<float> {
  using V4 = float4;
  // v_rsq_f32 (1 ulp).  Its argument is r^2 = d^2 + eps^2 of an interaction
  // that is not a self term; the input guard of rsqrtf() only matters for
  // denormal arguments, which would need |d| and eps below 1e-19.
  static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
  static __device__ __forceinline__ float fm(float a, float b, float c) { return fmaf(a, b, c); }
};
template <> struct Prec<double> {
  using V4 = double4;
  static __device__ __forceinline__ double rsq(double x) { return rsqrt(x); }
  static __device__ __forceinline__ double fm(double a, double b, double c) { return fma(a, b, c); }
};

template <typename R>
struct Acc {
  R s[4];  // running sums   ex, ey, ez, pot
  R c[4];  // compensation (FP32 only)
  R p[4];  // partial sums of the current batch
};

// M2P: entry = {cx,cy,cz,q | px,py,pz,2qx | 2qy,2qz,xy,yz | zx,tr,-,-}; the
// separation d is target minus centre of charge.  With w = 1/r^2, the matrix
// A2 = [[2qx,xy,zx],[xy,2qy,yz],[zx,yz,2qz]], g = A2 d, S2 = d.g, a = d.p + tr:
//   phi = (1/r) (q - w (a - 1.5 w S2))
//   E   = (w/r) (c d + p - 3 w g),   c = q - 3 w (a - 2.5 w S2)
// which is phi and -grad phi of module_coulomb_kernels.f90:60-145.
template <typename R>
__device__ __forceinline__ void m2p(R dx, R dy, R dz, const typename Prec<R>::V4& b,
                                    const typename Prec<R>::V4& c, const typename Prec<R>::V4& d,
                                    R q, R eps2, R* p) {
  using P = Prec<R>;
  const R r2 = P::fm(dz, dz, P::fm(dy, dy, P::fm(dx, dx, eps2)));
  const R rd = P::rsq(r2);
  const R w = rd * rd;
  const R rd3 = rd * w;
  const R px = b.x, py = b.y, pz = b.z, qx2 = b.w;
  const R qy2 = c.x, qz2 = c.y, xy = c.z, yz = c.w;
  const R zx = d.x, tr = d.y;
  const R gx = P::fm(qx2, dx, P::fm(xy, dy, zx * dz));
  const R gy = P::fm(qy2, dy, P::fm(xy, dx, yz * dz));
  const R gz = P::fm(qz2, dz, P::fm(yz, dy, zx * dx));
  const R s2 = P::fm(dz, gz, P::fm(dy, gy, dx * gx));
  const R av = P::fm(dz, pz, P::fm(dy, py, P::fm(dx, px, tr)));
  const R bv = w * s2;
  p[3] = P::fm(rd, P::fm(-w, P::fm(R(-1.5), bv, av), q), p[3]);
  const R w3 = R(3) * w;
  const R cc = P::fm(-w3, P::fm(R(-2.5), bv, av), q);
  p[0] = P::fm(rd3, P::fm(cc, dx, P::fm(-w3, gx, px)), p[0]);
  p[1] = P::fm(rd3, P::fm(cc, dy, P::fm(-w3, gy, py)), p[1]);
  p[2] = P::fm(rd3, P::fm(cc, dz, P::fm(-w3, gz, pz)), p[2]);
}

// P2P (module_coulomb_kernels.f90:242-257): phi = q/r, E = q d / r^3.
template <typename R>
__device__ __forceinline__ void p2p(R dx, R dy, R dz, R q, R eps2, bool self, R* p) {
  using P = Prec<R>;
  R r2 = P::fm(dz, dz, P::fm(dy, dy, P::fm(dx, dx, eps2)));
  r2 = self ? R(1) : r2;
  const R qe = self ? R(0) : q;
  const R rd = P::rsq(r2);
  const R qrd = qe * rd;
  const R qrd3 = qrd * (rd * rd);
  p[3] += qrd;
  p[0] = P::fm(qrd3, dx, p[0]);
  p[1] = P::fm(qrd3, dy, p[1]);
  p[2] = P::fm(qrd3, dz, p[2]);
}

template <typename R>
__device__ __forceinline__ void fold(Acc<R>& a) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if constexpr (sizeof(R) == 4) {   // Kahan step on the batch's partial sum
      const R y = a.p[k] - a.c[k];
      const R t = a.s[k] + y;
      a.c[k] = (t - a.s[k]) - y;
      a.s[k] = t;
    } else {
      a.s[k] += a.p[k];
    }
    a.p[k] = R(0);
  }
}

// Exact reference MAC for one target, FP64 without contraction.
__device__ __forceinline__ bool mac_fails(double xt, double yt, double zt, double cx, double cy,
                                          double cz, double b2, double theta2) {
#pragma clang fp contract(off)
  const double dx = xt - cx, dy = yt - cy, dz = zt - cz;
  const double d2 = dx * dx + dy * dy + dz * dz;
  return !(theta2 * d2 > b2);
}

// LM: capacity of the M2P list staged in LDS (the list tile); the P2P list
// holds kListP entries.  The default tiles (LM * sizeof(R) = 256 bytes per
// float component: FP32 64, FP64 32) leave LDS for four one-wave workgroups
// per SIMD; the register budget is held to the same four (at most 128 VGPRs).
// The larger tiles are LDS-limited to three.
template <typename R, int LM, bool PER, bool SINK>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(LM * sizeof(R) <= 256 ? 4 : 3)))
void walk_kernel(WalkArgs a) {
  using P = Prec<R>;
  using V4 = typename P::V4;
  // A step pops at most (cap - sp) / 7 nodes (each may push 8 children), so a
  // deep stack means small steps; wide near fields (small theta) fill it.
  // Four one-wave workgroups per SIMD leave 10 KB of LDS each: FP32 lists
  // take 5.4 KB, FP64 lists 6.4 KB, so only FP32 has room for the larger cap.
  constexpr int kCap = sizeof(R) == 4 ? kStackCap32 : kStackCap;
  __shared__ uint32_t stack[kCap + kStackSlack];
  __shared__ V4 lm[LM][4];
  __shared__ V4 lp[kListP];
  __shared__ uint32_t lpi[kListP];
  __shared__ uint32_t gsh;

  const int lane = threadIdx.x;
  const R eps2 = (R)a.eps2;
  const double theta2 = a.theta2;
  const uint32_t ngroups = SINK ? a.nsink_groups : a.sc->ngroups;
  const double* __restrict__ TX = SINK ? a.tx : a.x;
  const double* __restrict__ TY = SINK ? a.ty : a.y;
  const double* __restrict__ TZ = SINK ? a.tz : a.z;

  // Work queues, one per XCD: the groups (in key order) are cut into nqueue
  // contiguous ranges and workgroup b, which the dispatcher places on XCD
  // b mod nqueue, takes groups from range b mod nqueue first, so the waves
  // sharing an XCD's L2 walk neighbouring groups and reuse the same nodes.
  // An exhausted range sends the wave on to the next one.
  const uint32_t nq = (uint32_t)a.nqueue;
  uint32_t q = blockIdx.x % nq, qleft = nq;   // lane 0's position in the queues
  for (;;) {
    if (lane == 0) {
      uint32_t gg = ngroups;
      while (qleft) {
        const uint32_t lo = (uint32_t)((uint64_t)ngroups * q / nq);
        const uint32_t hi = (uint32_t)((uint64_t)ngroups * (q + 1) / nq);
        const uint32_t r = atomicAdd(a.queue + kQueueStride * q, 1u);
        if (r < hi - lo) { gg = lo + r; break; }
        q = q + 1 == nq ? 0u : q + 1;
        --qleft;
      }
      gsh = gg;
    }
    __syncthreads();
    const uint32_t g = gsh;
    __syncthreads();
    if (g >= ngroups) break;

    // Lanes s*cnt .. s*cnt+cnt-1 form slice s (s < S = 64/cnt) and hold the
    // group's targets in order; the 64 - S*cnt remaining lanes hold copies
    // of targets for the box and take no part in the sums.
    uint32_t a0, cnt;
    if (!SINK || a.gfirst) { a0 = a.gfirst[g]; cnt = a.gcount[g]; }
    else { a0 = g; cnt = 1u; }
    const uint32_t T = cnt;
    const uint32_t S = kWave / T;
    const uint32_t s = (uint32_t)lane / T;
    const uint32_t t = (uint32_t)lane - s * T;
    const bool valid = s < S;
    const uint32_t tid = a0 + t;
    const double xt = TX[tid], yt = TY[tid], zt = TZ[tid];
    const double ox = TX[a0], oy = TY[a0], oz = TZ[a0];
    const double lox = wave_min(xt), loy = wave_min(yt), loz = wave_min(zt);
    const double hix = wave_max(xt), hiy = wave_max(yt), hiz = wave_max(zt);
    // target position in the kernel's frame (FP64: set per box below)
    R tx = R(0), ty = R(0), tz = R(0);
    if constexpr (sizeof(R) == 4) {
      tx = (R)(xt - ox); ty = (R)(yt - oy); tz = (R)(zt - oz);
    }

    Acc<R> acc;
#pragma unroll
    for (int k = 0; k < 4; ++k) { acc.s[k] = R(0); acc.c[k] = R(0); acc.p[k] = R(0); }

    uint32_t sp = 0, nm = 0, np = 0, cnt_p2p = 0, cnt_m2p = 0, namb = 0, nsteps = 0;
    uint32_t mac_lane = 0;   // this lane's MAC evaluations, summed over the wave at the end
    bool overflow = false;
    const uint32_t root = (!SINK && a.sys) ? a.sysroot[a.sys[a0]] : 0u;   // the group's own root
    double sx = xt, sy = yt, sz = zt;    // targets of the current box: x - vbox

    // Batches: entry k goes to slice k mod S; the loop count is uniform.
    // FP32: two entries per pass (their LDS reads issued together), then the
    // tail; FP64 keeps one entry per pass (register pressure).
    auto flush_m2p = [&]() {
      uint32_t kk = 0;
      constexpr bool kPair = sizeof(R) == 4;
      const uint32_t n2 = (kPair && nm >= 2 * S) ? nm - 2 * S + 1 : 0;
      if (kPair && valid) {
        for (; kk < n2; kk += 2 * S) {
          const uint32_t k = kk + s, k2 = k + S;
          const V4 e0 = lm[k][0], e1 = lm[k][1], e2 = lm[k][2], e3 = lm[k][3];
          const V4 f0 = lm[k2][0], f1 = lm[k2][1], f2 = lm[k2][2], f3 = lm[k2][3];
          m2p<R>(tx - (R)e0.x, ty - (R)e0.y, tz - (R)e0.z, e1, e2, e3, (R)e0.w, eps2, acc.p);
          m2p<R>(tx - (R)f0.x, ty - (R)f0.y, tz - (R)f0.z, f1, f2, f3, (R)f0.w, eps2, acc.p);
        }
      }
      kk = n2 ? ((n2 - 1) / (2 * S) + 1) * 2 * S : 0;
      for (; kk < nm; kk += S) {
        const uint32_t k = kk + s;
        if (valid && k < nm) {
          const V4 e0 = lm[k][0], e1 = lm[k][1], e2 = lm[k][2], e3 = lm[k][3];
          m2p<R>(tx - (R)e0.x, ty - (R)e0.y, tz - (R)e0.z, e1, e2, e3, (R)e0.w, eps2, acc.p);
        }
      }
      fold(acc);
    };
    // self term: the reference's rule (a zero separation from the leaf's
    // particle) for sinks, the particle index for the grown particles
    auto p2p_self = [&](uint32_t k) -> bool {
      if constexpr (SINK) {
        const uint32_t j = lpi[k];
        return sx == a.x[j] && sy == a.y[j] && sz == a.z[j];
      } else {
        return lpi[k] == tid;   // mirror-box entries hold ~0u
      }
    };
    // FP32: two entries per pass as for M2P (same per-lane order of terms)
    auto flush_p2p = [&]() {
      uint32_t kk = 0;
      constexpr bool kPair = sizeof(R) == 4;
      const uint32_t n2 = (kPair && np >= 2 * S) ? np - 2 * S + 1 : 0;
      if (kPair && valid) {
        for (; kk < n2; kk += 2 * S) {
          const uint32_t k = kk + s, k2 = k + S;
          const V4 e = lp[k], f = lp[k2];
          const bool self1 = p2p_self(k), self2 = p2p_self(k2);
          p2p<R>(tx - (R)e.x, ty - (R)e.y, tz - (R)e.z, (R)e.w, eps2, self1, acc.p);
          p2p<R>(tx - (R)f.x, ty - (R)f.y, tz - (R)f.z, (R)f.w, eps2, self2, acc.p);
        }
      }
      kk = n2 ? ((n2 - 1) / (2 * S) + 1) * 2 * S : 0;
      for (; kk < np; kk += S) {
        const uint32_t k = kk + s;
        if (valid && k < np) {
          const V4 e = lp[k];
          p2p<R>(tx - (R)e.x, ty - (R)e.y, tz - (R)e.z, (R)e.w, eps2, p2p_self(k), acc.p);
        }
      }
      fold(acc);
    };

    const int nbox = PER ? a.nbox : 1;
    for (int box = 0; box < nbox && !overflow; ++box) {
    double vx = 0.0, vy = 0.0, vz = 0.0;
    if constexpr (PER) { vx = a.vbox[3 * box]; vy = a.vbox[3 * box + 1]; vz = a.vbox[3 * box + 2]; }
    const bool central = !PER || (vx == 0.0 && vy == 0.0 && vz == 0.0);
    // targets of this box: x - vbox; group box likewise (rounding is monotone,
    // so the shifted extremes are the extremes of the shifted targets)
    sx = xt - vx; sy = yt - vy; sz = zt - vz;
    const double blox = lox - vx, bloy = loy - vy, bloz = loz - vz;
    const double bhix = hix - vx, bhiy = hiy - vy, bhiz = hiz - vz;
    const double fox = ox - vx, foy = oy - vy, foz = oz - vz;   // FP32 node frame
    if constexpr (sizeof(R) == 8) { tx = sx; ty = sy; tz = sz; }
    if (lane == 0) stack[0] = root | kTwigBit;
    sp = 1;
    __syncthreads();
    while (sp > 0) {
      // Up to min(64, LM) nodes are read from the top of the stack (so one
      // step's appends always fit an emptied list, positions nm + rank < LM);
      // after the box test only the longest run of them whose pushes fit
      // below the cap is visited, the rest stay on the stack.
      constexpr uint32_t kPop = (uint32_t)(LM < kWave ? LM : kWave);
      static_assert(kListP >= kWave, "P2P list must hold one full step");
      const uint32_t m = sp < kPop ? sp : kPop;
      if (nm + m > (uint32_t)LM) { __syncthreads(); flush_m2p(); nm = 0; }
      if (np + m > (uint32_t)kListP) { __syncthreads(); flush_p2p(); np = 0; }
      const bool has = (uint32_t)lane < m;
      const uint32_t entry = has ? stack[sp - 1 - lane] : 0u;
      const uint32_t node = entry & ~kTwigBit;
      const bool twig_entry = (entry & kTwigBit) != 0u;
      __syncthreads();
      ++nsteps;

      // All of a node's data is requested at once, before any decision, so a
      // traversal step waits for one memory round trip.  A twig also brings
      // its walk moments and its children's kinds; a leaf has neither.
      NodeGeom gm = {0.0, 0.0, 0.0, 0.0};
      NodeLink lk = {0u, 0u, 0u, 0u};
      V4 w0, w1, w2;
      uint32_t kids = 0u;
      if (has) {
        gm = a.geom[node];
        lk = a.link[node];
      }
      if (has && twig_entry) {
        kids = a.kidtwig[node];
        const V4* src;
        if constexpr (sizeof(R) == 4) src = reinterpret_cast<const V4*>(&a.wm32[node]);
        else src = reinterpret_cast<const V4*>(&a.wm64[node]);
        w0 = src[0]; w1 = src[1]; w2 = src[2];
      }
      // Decide from the group's bounding box when possible (margins cover
      // every rounding of the exact per-target test below).
      bool acc_ok = false, rej = false;
      uint32_t macc = 0;
      if (has && lk.nchild != 0u) {
        const double ex = fmax(fmax(blox - gm.cx, gm.cx - bhix), 0.0);
        const double ey = fmax(fmax(bloy - gm.cy, gm.cy - bhiy), 0.0);
        const double ez = fmax(fmax(bloz - gm.cz, gm.cz - bhiz), 0.0);
        const double fx = fmax(fabs(gm.cx - blox), fabs(gm.cx - bhix));
        const double fy = fmax(fabs(gm.cy - bloy), fabs(gm.cy - bhiy));
        const double fz = fmax(fabs(gm.cz - bloz), fabs(gm.cz - bhiz));
        const double dmin2 = ex * ex + ey * ey + ez * ez;
        const double dmax2 = fx * fx + fy * fy + fz * fz;
        acc_ok = (theta2 * dmin2) * (1.0 - 1e-12) > gm.b2;
        rej = !acc_ok && (theta2 * dmax2) * (1.0 + 1e-12) <= gm.b2;
        macc = acc_ok ? cnt : (rej ? 1u : 0u);
      }
      // Visit lanes 0 .. k-1 for the largest k whose step leaves at most kCap
      // entries: sp - k + (children of lanes < k, counted for every twig the
      // box did not accept) <= kCap.  k >= 1 always, so the stack passes the
      // cap only by the one-node steps of a depth-first descent (within
      // kStackSlack), and k is never below the worst-case count
      // (kCap - sp) / 7.  Nodes k .. m-1 stay on the stack unvisited: none of
      // their decisions is counted.
      // While the worst case fits (7 net pushes per node), all m are visited
      // without the search.
      uint32_t k = m;
      if ((int)(7u * m) > (int)kCap - (int)sp) {
        const uint32_t pmax = (has && lk.nchild != 0u && !acc_ok) ? lk.nchild : 0u;
        uint32_t ptot;
        const uint32_t pex = wave_exclusive_scan16(pmax, &ptot);
        const bool fits = has && (int)(pex + pmax) - (lane + 1) <= (int)kCap - (int)sp;
        const unsigned long long fm = __ballot(fits);
        k = fm ? 64u - (uint32_t)__clzll((long long)fm) : 1u;
      }
      const bool keep = (uint32_t)lane < k;
      sp -= k;
      const bool leaf = keep && lk.nchild == 0u;
      const bool twig = keep && lk.nchild != 0u;
      // Undecided nodes: exact per-target test.  Every slice holds all the
      // targets, so slice s tests the s-th undecided node of the batch and S
      // nodes are settled per pass.  The nodes' centres and b^2 are staged in
      // the free tail of the M2P list (entries nm .. nm + m - 1, which this
      // step's appends overwrite only after the tests), so each slice reads
      // its node with one broadcast load per pass.
      const bool amb_me = twig && !acc_ok && !rej;
      const unsigned long long amb = __ballot(amb_me);
      const uint32_t na = (uint32_t)popc64(amb);
      namb += na;
      if (na) {
        const uint32_t my_rank = mask_rank(amb);
        if (amb_me) {
          double2* st = reinterpret_cast<double2*>(&lm[nm + my_rank][0]);
          st[0] = make_double2(gm.cx, gm.cy);
          st[1] = make_double2(gm.cz, gm.b2);
        }
        __syncthreads();
        for (uint32_t base = 0; base < na; base += S) {
          const uint32_t r = base + s;
          const bool active = valid && r < na;
          double2 c0 = make_double2(0.0, 0.0), c1 = c0;
          if (active) {
            const double2* st = reinterpret_cast<const double2*>(&lm[nm + r][0]);
            c0 = st[0];
            c1 = st[1];
          }
          const bool fail = active && mac_fails(sx, sy, sz, c0.x, c0.y, c1.x, c1.y, theta2);
          const unsigned long long ball = __ballot(fail);
          if (amb_me && my_rank >= base && my_rank < base + S) {
            const uint32_t rr = my_rank - base;
            const unsigned long long fm =
                (ball >> (rr * T)) & (T == 64 ? ~0ull : ((1ull << T) - 1ull));
            if (fm == 0ull) { acc_ok = true; macc = cnt; }
            else { rej = true; macc = (uint32_t)__ffsll((long long)fm); }
          }
        }
        __syncthreads();
      }
      mac_lane += twig ? macc : 0u;

      const bool take_m = twig && acc_ok;
      const unsigned long long am = __ballot(take_m);
      const unsigned long long pm = __ballot(leaf);
      if (take_m) {
        const uint32_t pos = nm + mask_rank(am);
        V4 e0;
        if constexpr (sizeof(R) == 4)
          e0 = make_float4((float)(gm.cx - fox), (float)(gm.cy - foy), (float)(gm.cz - foz), 0.0f);
        else
          e0 = make_double4(gm.cx, gm.cy, gm.cz, 0.0);
        // w0 = {q, px, py, pz}, w1 = {2qx, 2qy, 2qz, xy}, w2 = {yz, zx, tr, -}
        e0.w = w0.x;
        V4 e1, e2, e3;
        e1.x = w0.y; e1.y = w0.z; e1.z = w0.w; e1.w = w1.x;
        e2.x = w1.y; e2.y = w1.z; e2.z = w1.w; e2.w = w2.x;
        e3.x = w2.y; e3.y = w2.z; e3.z = 0; e3.w = 0;
        lm[pos][0] = e0; lm[pos][1] = e1; lm[pos][2] = e2; lm[pos][3] = e3;
      }
      if (leaf) {
        const uint32_t pos = np + mask_rank(pm);
        V4 e;
        if constexpr (sizeof(R) == 4) {
          e = make_float4((float)(gm.cx - fox), (float)(gm.cy - foy), (float)(gm.cz - foz),
                          (float)gm.b2);
        } else {
          e = make_double4(gm.cx, gm.cy, gm.cz, gm.b2);
        }
        lp[pos] = e;
        lpi[pos] = (SINK || central) ? lk.first : ~0u;
      }
      const uint32_t nma = (uint32_t)popc64(am), npa = (uint32_t)popc64(pm);
      nm += nma; np += npa; cnt_m2p += nma; cnt_p2p += npa;

      const uint32_t nc = (twig && !acc_ok) ? lk.nchild : 0u;   // at most 8
      uint32_t tot;
      const uint32_t excl = wave_exclusive_scan16(nc, &tot);
      if (sp + tot > (uint32_t)(kCap + kStackSlack)) { overflow = true; break; }
      for (uint32_t k = 0; k < nc; ++k)
        stack[sp + excl + k] = (lk.first + k) | (((kids >> k) & 1u) ? kTwigBit : 0u);
      sp += tot;
      __syncthreads();
    }
    if constexpr (PER && (sizeof(R) == 8 || SINK)) {   // next box shifts the targets
      __syncthreads();
      if (nm) flush_m2p();
      if (np) flush_p2p();
      nm = 0; np = 0;
    }
    }   // boxes
    __syncthreads();
    if (nm) flush_m2p();
    if (np) flush_p2p();

    // combine slices (target t lives in lanes t, t+T, t+2T, ...) in FP64
    double out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double v0 = (double)acc.s[k] - (double)acc.c[k];
      double v = v0;
      for (uint32_t o = 1; o < S; ++o) v += __shfl(v0, (int)(t + o * T));
      out[k] = v * a.sign;
    }
    if (overflow) {
      if (lane == 0) atomicOr(&a.sc->flags, kFlagStackOverflow);
    }
    if (s == 0 && valid) {
      a.ex[tid] = out[0];
      a.ey[tid] = out[1];
      a.ez[tid] = out[2];
      a.pot[tid] = out[3];
      a.work[tid] = (SINK ? 0.0 : 1.0) + (double)(cnt_p2p + cnt_m2p);
      a.np2p[tid] = cnt_p2p;
      a.nm2p[tid] = cnt_m2p;
    }
    const unsigned long long nmac = SINK ? 0ull : wave_sum((unsigned long long)mac_lane);
    if (!SINK && lane == 0) {   // per-group record, summed by walk_stats_kernel
      uint4* gs = reinterpret_cast<uint4*>(a.gstat + 8 * (size_t)g);
      gs[0] = make_uint4((uint32_t)nmac, cnt_p2p, cnt_m2p, namb);
      gs[1] = make_uint4(nsteps, cnt, 0u, 0u);
    }
    __syncthreads();
  }
}

// Sums the per-group records into the device scalars (walk statistics of
// tree_walk_statistics, module_tree_walk_simple.f90:63-115).
__global__ __launch_bounds__(256) void walk_stats_kernel(const uint32_t* __restrict__ gstat,
                                                        DevScalars* __restrict__ sc) {
  const int64_t ng = sc->ngroups;
  unsigned long long v[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng;
       g += (int64_t)gridDim.x * blockDim.x) {
    const uint4 r0 = reinterpret_cast<const uint4*>(gstat + 8 * g)[0];
    const uint4 r1 = reinterpret_cast<const uint4*>(gstat + 8 * g)[1];
    const unsigned long long c = r1.y;
    v[0] += c * (r0.y + r0.z);
    v[1] += r0.x;
    v[2] += c * r0.y;
    v[3] += c * r0.z;
    v[4] += r0.w;
    v[5] += r1.x;
  }
  __shared__ unsigned long long red[6][4];
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const unsigned long long s = wave_sum(v[k]);
    if (lane_id() == 0) red[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    const unsigned long long s = red[k][0] + red[k][1] + red[k][2] + red[k][3];
    if (s) {
      unsigned long long* dst[6] = {&sc->interactions, &sc->mac_evals, &sc->p2p, &sc->m2p,
                                    &sc->amb_nodes, &sc->walk_steps};
      atomicAdd(dst[k], s);
    }
  }
}

__global__ void walk_reset_kernel(DevScalars* sc, uint32_t* queue) {
  if (threadIdx.x < kMaxQueues) queue[kQueueStride * threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    sc->interactions = 0;
    sc->mac_evals = 0;
    sc->p2p = 0;
    sc->m2p = 0;
    sc->amb_nodes = 0;
    sc->walk_steps = 0;
  }
}

__global__ void queue_reset_kernel(uint32_t* queue) {
  if (threadIdx.x < kMaxQueues) queue[kQueueStride * threadIdx.x] = 0;
}

// FP64 direct sum (src/utils/module_directsum.f90:34-172): all sources
// against a list of targets, self excluded by index, reference P2P form.
__global__ __launch_bounds__(256) void direct_kernel(int64_t n, const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z,
                                                    const double* __restrict__ q, int64_t ntest,
                                                    const int64_t* __restrict__ testidx, double eps2,
                                                    double sign, double* __restrict__ ex,
                                                    double* __restrict__ ey, double* __restrict__ ez,
                                                    double* __restrict__ pot, int nbox,
                                                    const double* __restrict__ vbox) {
#pragma clang fp contract(off)
  __shared__ double sx[256], sy[256], sz[256], sq[256];
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t tj = j < ntest ? testidx[j] : -1;
  const double xt = tj >= 0 ? x[tj] : 0.0, yt = tj >= 0 ? y[tj] : 0.0, zt = tj >= 0 ? z[tj] : 0.0;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t base = 0; base < n; base += 256) {
    const int64_t i = base + threadIdx.x;
    __syncthreads();
    sx[threadIdx.x] = i < n ? x[i] : 0.0;
    sy[threadIdx.x] = i < n ? y[i] : 0.0;
    sz[threadIdx.x] = i < n ? z[i] : 0.0;
    sq[threadIdx.x] = i < n ? q[i] : 0.0;
    __syncthreads();
    const int lim = (n - base) < 256 ? (int)(n - base) : 256;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    if (!vbox) {
      for (int k = 0; k < lim; ++k) {
        if (base + k == tj) continue;
        const double dx = xt - sx[k], dy = yt - sy[k], dz = zt - sz[k];
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(d2 + eps2);
        const double rd = 1.0 / r;
        const double rd3q = sq[k] * rd * rd * rd;
        part[3] += sq[k] * rd;
        part[0] += rd3q * dx;
        part[1] += rd3q * dy;
        part[2] += rd3q * dz;
      }
    } else {
      // every source in every mirror box (the target shifted by -v), self
      // only in the central box (module_directsum.f90:80-93)
      for (int k = 0; k < lim; ++k) {
        for (int b = 0; b < nbox; ++b) {
          const double vx = vbox[3 * b], vy = vbox[3 * b + 1], vz = vbox[3 * b + 2];
          if (base + k == tj && vx == 0.0 && vy == 0.0 && vz == 0.0) continue;
          const double dx = (xt - vx) - sx[k], dy = (yt - vy) - sy[k], dz = (zt - vz) - sz[k];
          const double d2 = dx * dx + dy * dy + dz * dz;
          const double r = sqrt(d2 + eps2);
          const double rd = 1.0 / r;
          const double rd3q = sq[k] * rd * rd * rd;
          part[3] += sq[k] * rd;
          part[0] += rd3q * dx;
          part[1] += rd3q * dy;
          part[2] += rd3q * dz;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] += part[c];
  }
  if (j < ntest) {
    ex[j] = sign * acc[0];
    ey[j] = sign * acc[1];
    ez[j] = sign * acc[2];
    pot[j] = sign * acc[3];
  }
}

}  // namespace

// Persistent grid: every wavefront the device can hold for this instance.
int walk_default_blocks(int precision, int list_tile) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
  }
  int per = 0;
  hipError_t e;
  if (precision == kWalkFP64)
    e = list_tile == 64
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, walk_kernel<double, 64, false, false>, kWave, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, walk_kernel<double, 32, false, false>, kWave, 0);
  else
    e = list_tile == 128
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, walk_kernel<float, 128, false, false>, kWave, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, walk_kernel<float, 64, false, false>, kWave, 0);
  if (e != hipSuccess || per < 1) per = 8;
  return cus * per;
}

void launch_walk_reset(hipStream_t s, const WalkArgs& a) {
  walk_reset_kernel<<<1, 64, 0, s>>>(a.sc, a.queue);
}

void launch_queue_reset(hipStream_t s, uint32_t* queue) { queue_reset_kernel<<<1, 64, 0, s>>>(queue); }

bool walk_list_tile_ok(int precision, int list_tile) {
  if (list_tile == 0) return true;
  return precision == kWalkFP64 ? (list_tile == 32 || list_tile == 64)
                                : (list_tile == 64 || list_tile == 128);
}

template <bool PER, bool SINK>
static void launch_walk_t(hipStream_t s, const WalkArgs& a, int precision, int blocks,
                          int list_tile) {
  // default (0): the smaller tile of each precision, whose lower LDS use
  // lets one more wavefront per SIMD resident (measured faster on MI355X)
  if (precision == kWalkFP64) {
    if (list_tile == 64)
      walk_kernel<double, 64, PER, SINK><<<blocks, kWave, 0, s>>>(a);
    else
      walk_kernel<double, 32, PER, SINK><<<blocks, kWave, 0, s>>>(a);
  } else {
    if (list_tile == 128)
      walk_kernel<float, 128, PER, SINK><<<blocks, kWave, 0, s>>>(a);
    else
      walk_kernel<float, 64, PER, SINK><<<blocks, kWave, 0, s>>>(a);
  }
}

void launch_walk(hipStream_t s, const WalkArgs& a, int precision, int blocks, int list_tile) {
  const bool sink = a.tx != nullptr;
  if (a.nbox > 1)
    sink ? launch_walk_t<true, true>(s, a, precision, blocks, list_tile)
         : launch_walk_t<true, false>(s, a, precision, blocks, list_tile);
  else
    sink ? launch_walk_t<false, true>(s, a, precision, blocks, list_tile)
         : launch_walk_t<false, false>(s, a, precision, blocks, list_tile);
}

void launch_walk_stats(hipStream_t s, const WalkArgs& a) {
  walk_stats_kernel<<<256, 256, 0, s>>>(a.gstat, a.sc);
}

void launch_direct(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                   const double* q, int64_t ntest, const int64_t* testidx, double eps2, double sign,
                   double* ex, double* ey, double* ez, double* pot, int nbox, const double* vbox) {
  if (ntest <= 0) return;
  direct_kernel<<<(unsigned)((ntest + 255) / 256), 256, 0, s>>>(n, x, y, z, q, ntest, testidx, eps2,
                                                                sign, ex, ey, ez, pot, nbox, vbox);
}

}  // namespace pepc
