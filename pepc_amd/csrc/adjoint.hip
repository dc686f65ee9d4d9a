// adjoint.hip -- gradient of the tree solve with respect to particle
// positions and charges (no reference counterpart: PEPC computes no
// derivatives).
//
// With the tree topology and the interaction lists held fixed ("frozen"),
// the fields are smooth functions of the particles.  For a loss
// L = sum_i gpot_i pot_i + gE_i . E_i, every interaction of target i with a
// node (centre c, moments q, p, Qt) contributes, with d = x_i - c:
//   dL/dd (to the target, and minus it to the centre),
//   dL/dq, dL/dp, dL/dQt (to the node's moments),
// where phi = q G + p.grad G + Qt : grad grad G, E = -grad phi and
// G = (d^2 + eps^2)^(-1/2) (the polynomial of module_coulomb_kernels.f90:
// 60-145).  A leaf (P2P) is the case p = Qt = 0 with the source particle in
// place of the node.  Node moments are closed forms over the node's
// particles (q = sum q_j, p = sum q_j s_j, Qt = 1/2 sum q_j s_j s_j^T,
// s_j = c - x_j, c = sum |q_j| x_j / sum |q_j|), which is what the upward pass
// computes (shift_multipoles_up, coulomb/module_interaction_specific.f90:
// 84-140); the node adjoints therefore go to each particle below the node
// directly.
//
// Kernels: (1) a group walk with the reference acceptance rule (a node is
// accepted when every target of the group passes the test, as
// module_tree_walk_simple.f90:301-384) that evaluates the per-interaction
// derivatives for the group's targets, reduces them across the wavefront
// and adds them to the node and source-particle adjoints with FP64 atomics;
// (2) the centre adjoint of every twig; (3) one thread per particle climbing
// its ancestors.  FP64 throughout.
#include "kernels.h"

namespace pepc {

namespace {

constexpr int kT = 256;
constexpr int kStackCap = 512;
constexpr int kStackSlack = 192;   // >= 7 * 21 + 8, as the forward walk

__device__ __forceinline__ bool broken(const DevScalars* sc) {
  return (sc->flags & (kFlagNodeOverflow | kFlagDupKeys | kFlagBadBox)) != 0u;
}

// Derivatives of gphi * phi + u . E for one interaction (see header comment).
// qt = {qx, qy, qz, xy, yz, zx} (Qt_xx = qx, Qt_xy = xy / 2, ...).  Outputs:
// D = dL/dd, gq, gp, M = dL/dQt as the symmetric matrix {xx, yy, zz, xy, yz, zx}.
// The adjoint is checked against the oracle to a tolerance, so its terms may
// contract products and sums into fused multiply-adds.
__device__ void adjoint_term(const double d[3], double r2, double gphi, const double u[3],
                             double q, const double p[3], const double qt[6], double D[3],
                             double& gq, double gp[3], double M[6]) {
#pragma clang fp contract(fast)
  const double ri = rsqrt(r2);
  const double ri2 = ri * ri;
  const double r3 = ri * ri2, r5 = r3 * ri2, r7 = r5 * ri2, r9 = r7 * ri2;
  const double ud = u[0] * d[0] + u[1] * d[1] + u[2] * d[2];
  const double pd = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
  const double up = u[0] * p[0] + u[1] * p[1] + u[2] * p[2];
  const double hxy = 0.5 * qt[3], hyz = 0.5 * qt[4], hzx = 0.5 * qt[5];
  const double Qd[3] = {qt[0] * d[0] + hxy * d[1] + hzx * d[2],
                        hxy * d[0] + qt[1] * d[1] + hyz * d[2],
                        hzx * d[0] + hyz * d[1] + qt[2] * d[2]};
  const double Qu[3] = {qt[0] * u[0] + hxy * u[1] + hzx * u[2],
                        hxy * u[0] + qt[1] * u[1] + hyz * u[2],
                        hzx * u[0] + hyz * u[1] + qt[2] * u[2]};
  const double tr = qt[0] + qt[1] + qt[2];
  const double dQd = d[0] * Qd[0] + d[1] * Qd[1] + d[2] * Qd[2];
  const double uQd = u[0] * Qd[0] + u[1] * Qd[1] + u[2] * Qd[2];
  // The terms of phi = q G + p.grad G + Qt : grad grad G and E = -grad phi,
  // differentiated and collected by vector, with
  //   s1 = gphi r^-3 + 3 ud r^-5,  s2 = 3 gphi r^-5 + 15 ud r^-7:
  //   gp = u r^-3 - s1 d,  M_ab = s2 d_a d_b - 3 r^-5 (u_a d_b + u_b d_a) - s1 delta_ab,
  //   D = cu u - s1 p + 2 s2 Qd - 6 r^-5 Qu + cd d.
  const double s1 = gphi * r3 + 3.0 * ud * r5;
  const double s2 = 3.0 * gphi * r5 + 15.0 * ud * r7;
  const double s3 = 15.0 * gphi * r7 + 105.0 * ud * r9;
  const double pt = pd + tr;
  const double t5 = 3.0 * r5;
  gq = gphi * ri + ud * r3;
#pragma unroll
  for (int k = 0; k < 3; ++k) gp[k] = u[k] * r3 - s1 * d[k];
  const int ma[6] = {0, 1, 2, 0, 1, 2}, mb[6] = {0, 1, 2, 1, 2, 0};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int a = ma[k], b = mb[k];
    const double v = s2 * d[a] * d[b] - t5 * (u[a] * d[b] + u[b] * d[a]);
    M[k] = a == b ? v - s1 : v;
  }
  const double cu = q * r3 - t5 * pt + 15.0 * r7 * dQd;
  const double cd = -q * s1 + pt * s2 - t5 * up - s3 * dQd + 30.0 * r7 * uQd;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    D[k] = cu * u[k] - s1 * p[k] + 2.0 * s2 * Qd[k] - 2.0 * t5 * Qu[k] + cd * d[k];
}

// The same for a leaf (p = Qt = 0): D = dL/dd and gq only.
__device__ __forceinline__ void adjoint_term_p2p(const double d[3], double r2, double gphi,
                                                 const double u[3], double q, double D[3],
                                                 double& gq) {
#pragma clang fp contract(fast)
  const double ri = rsqrt(r2);
  const double ri2 = ri * ri;
  const double r3 = ri * ri2, r5 = r3 * ri2;
  const double ud = u[0] * d[0] + u[1] * d[1] + u[2] * d[2];
  gq = gphi * ri + ud * r3;
#pragma unroll
  for (int k = 0; k < 3; ++k) D[k] = q * ((u[k] - gphi * d[k]) * r3 - 3.0 * ud * d[k] * r5);
}

// Lane k's value in every lane (k uniform): register reads, no LDS.
__device__ __forceinline__ uint32_t lane_bcast(uint32_t v, int k) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
}
__device__ __forceinline__ double lane_bcast(double v, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}

// v + (v of the lane that the DPP control selects), 64-bit as two halves.
template <int kCtrl>
__device__ __forceinline__ double dpp_add(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, false);
  return v + __hiloint2double(hi, lo);
}

// Exchanges between two registers, 64-bit as two halves:
//   swap32: x = [x lanes 0-31, y lanes 0-31], y = [x lanes 32-63, y lanes 32-63];
//   swap16: x = [x row 0, y row 0, x row 2, y row 2], y = [x row 1, y row 1,
//           x row 3, y row 3] (rows of 16 lanes).
__device__ __forceinline__ void swap32(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x),
                                                   (unsigned)__double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x),
                                                   (unsigned)__double2hiint(y), false, false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ void swap16(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x),
                                                   (unsigned)__double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x),
                                                   (unsigned)__double2hiint(y), false, false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}

// Sum over the 16 lanes of each row, in every lane of the row.
__device__ __forceinline__ double row_sum(double v) {
  v = dpp_add<0xB1>(v);    // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);    // quad_perm [2,3,0,1]
  v = dpp_add<0x124>(v);   // row_ror:4
  return dpp_add<0x128>(v);   // row_ror:8
}

// Wave sums of four values at once: the 32-lane halves of values 0 and 1
// (and of 2 and 3) are exchanged and added, then the rows of the two
// results, then each row is summed by DPP.  Row r of the result holds the sum
// of value row_value(r) = 0, 2, 1, 3 for r = 0..3; every lane of the row has
// it.
__device__ __forceinline__ int row_value(int r) { return ((r & 1) << 1) | (r >> 1); }
__device__ __forceinline__ double wsum4(double v0, double v1, double v2, double v3) {
  swap32(v0, v1);
  swap32(v2, v3);
  double x = v0 + v1, y = v2 + v3;
  swap16(x, y);
  return row_sum(x + y);
}

// The adjoint buffers are plain device allocations (coarse grained), where the
// hardware FP64 add atomic is exact.

__global__ void adjoint_clear_kernel(int64_t n, int64_t cap, double* __restrict__ gx,
                                     double* __restrict__ gy, double* __restrict__ gz,
                                     double* __restrict__ gq, double* __restrict__ nadj,
                                     uint32_t* __restrict__ queue, const DevScalars* __restrict__ sc) {
  const int64_t nn = (int64_t)sc->nnodes < cap ? (int64_t)sc->nnodes : cap;
  const int64_t m = broken(sc) ? 0 : nn * 16;
  if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n || i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n) { gx[i] = 0.0; gy[i] = 0.0; gz[i] = 0.0; gq[i] = 0.0; }
    if (i < m) nadj[i] = 0.0;
  }
}

// Group traversal with the reference acceptance rule (module_tree_walk_simple.f90:
// 301-384): a twig is accepted when every target of the group passes
// theta2 d^2 > b^2(level) (exact FP64, no contraction), else its children are
// visited; a leaf is always an interaction.  on_entry(node, first, nchild,
// cx, cy, cz) runs, uniformly across the wave, for every leaf and accepted
// twig.  The pop policy is the forward walk's (at most 64 nodes per step,
// room for 8 children each), so the stack never overflows; returns true if
// it would have.
template <typename OnEntry>
__device__ bool traverse_group(uint32_t root, double xt, double yt, double zt, bool valid,
                               double theta2, const NodeLink* __restrict__ link,
                               const NodeGeom* __restrict__ geom, uint32_t* stack,
                               OnEntry&& on_entry) {
  const int lane = threadIdx.x;
  uint32_t sp = 1;
  if (lane == 0) stack[0] = root;
  __syncthreads();
  while (sp > 0) {
    int room = ((int)kStackCap - (int)sp) / 7;
    room = room < 1 ? 1 : room;
    uint32_t m = sp < (uint32_t)room ? sp : (uint32_t)room;
    m = m > (uint32_t)kWave ? (uint32_t)kWave : m;
    const bool has = (uint32_t)lane < m;
    const uint32_t mynode = has ? stack[sp - 1 - lane] : 0u;
    NodeLink mylk = {0u, 0u, 0u, 0u};
    NodeGeom mygm = {0.0, 0.0, 0.0, 0.0};
    if (has) { mylk = link[mynode]; mygm = geom[mynode]; }
    __syncthreads();
    sp -= m;
    for (uint32_t k = 0; k < m; ++k) {
      const uint32_t node = lane_bcast(mynode, (int)k);
      const uint32_t first = lane_bcast(mylk.first, (int)k);
      const uint32_t nchild = lane_bcast(mylk.nchild, (int)k);
      const double cx = lane_bcast(mygm.cx, (int)k), cy = lane_bcast(mygm.cy, (int)k),
                   cz = lane_bcast(mygm.cz, (int)k), b2 = lane_bcast(mygm.b2, (int)k);
      if (nchild != 0u) {
        bool fail;
        {
#pragma clang fp contract(off)
          const double dx = xt - cx, dy = yt - cy, dz = zt - cz;
          const double d2 = dx * dx + dy * dy + dz * dz;
          fail = valid && !(theta2 * d2 > b2);
        }
        if (__ballot(fail) != 0ull) {
          if (sp + nchild > (uint32_t)(kStackCap + kStackSlack)) return true;
          if ((uint32_t)lane < nchild) stack[sp + lane] = first + (uint32_t)lane;
          sp += nchild;
          continue;
        }
      }
      on_entry(node, first, nchild, cx, cy, cz);
    }
    __syncthreads();
  }
  return false;
}

struct AdjArgs {
  const double *x, *y, *z, *q;
  const NodeGeom* geom;
  const NodeLink* link;
  const double* mom64;
  const uint32_t *gfirst, *gcount;
  const uint32_t *sys, *sysroot;
  DevScalars* sc;
  double theta2, eps2, sign;
  const double *gex, *gey, *gez, *gpot;
  double *gx, *gy, *gz, *gq;
  double* nadj;
  uint32_t* queue;
};

// The traversal is traverse_group's, written out so that each step also
// stages the popped nodes' interaction data in LDS with the node records (a
// twig's charge, dipole and quadrupole, a leaf's particle charge): the
// entries of a step then read LDS only, and the step's one wait on memory
// also drains the previous step's atomics.  A node's 13 sums (4 for a leaf's
// particle) are reduce-scattered across the wave (wsum4) and go out as one
// atomic instruction, one value per lane.  Four waves per SIMD: the 128-VGPR
// budget costs a few bytes of spill and hides more of each step's latency.
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4)))
void adjoint_walk_kernel(AdjArgs a) {
  __shared__ uint32_t stack[kStackCap + kStackSlack];
  __shared__ double smom[kWave][10];   // q, p[3], Qt[6] of popped node k
  __shared__ uint32_t gsh;
  const int lane = threadIdx.x;
  const uint32_t ngroups = a.sc->ngroups;
  const double eps2 = a.eps2;
  // Work units: a group's walk is cut at its root into one unit per child of
  // the root (the acceptance decisions do not depend on the order of the
  // walk, so the units' entries are the group's), which evens out the end of
  // the launch; unit 0 takes the root itself when the root is an interaction.
  const uint32_t split = ngroups <= (1u << 28) ? 8u : 1u;
  const uint32_t nunits = ngroups * split;
  for (;;) {
    if (lane == 0) gsh = atomicAdd(a.queue, 1u);
    __syncthreads();
    const uint32_t unit = gsh;
    __syncthreads();
    if (unit >= nunits) break;
    const uint32_t g = unit / split, part = unit - g * split;
    const uint32_t a0 = a.gfirst[g], cnt = a.gcount[g];
    const bool valid = (uint32_t)lane < cnt;
    const uint32_t tid = a0 + (valid ? (uint32_t)lane : 0u);
    const double xt = a.x[tid], yt = a.y[tid], zt = a.z[tid];
    const uint32_t root = a.sys ? a.sysroot[a.sys[a0]] : 0u;
    uint32_t start = root;
    if (split > 1) {
      const NodeLink rl = a.link[root];
      bool open = false;
      if (rl.nchild != 0u) {
        const NodeGeom rg = a.geom[root];
        bool fail;
        {
#pragma clang fp contract(off)
          const double dx = xt - rg.cx, dy = yt - rg.cy, dz = zt - rg.cz;
          const double d2 = dx * dx + dy * dy + dz * dz;
          fail = valid && !(a.theta2 * d2 > rg.b2);
        }
        open = __ballot(fail) != 0ull;
      }
      if (open ? part >= rl.nchild : part != 0u) continue;   // no such unit
      if (open) start = rl.first + part;
    }
    const double gphi = valid ? a.sign * a.gpot[tid] : 0.0;
    const double u[3] = {valid ? a.sign * a.gex[tid] : 0.0, valid ? a.sign * a.gey[tid] : 0.0,
                         valid ? a.sign * a.gez[tid] : 0.0};
    double tD[3] = {0.0, 0.0, 0.0};
    bool overflow = false;
    uint32_t sp = 1;
    if (lane == 0) stack[0] = start;
    __syncthreads();
    while (sp > 0 && !overflow) {
      int room = ((int)kStackCap - (int)sp) / 7;
      room = room < 1 ? 1 : room;
      uint32_t m = sp < (uint32_t)room ? sp : (uint32_t)room;
      m = m > (uint32_t)kWave ? (uint32_t)kWave : m;
      const bool has = (uint32_t)lane < m;
      const uint32_t mynode = has ? stack[sp - 1 - lane] : 0u;
      NodeLink mylk = {0u, 0u, 0u, 0u};
      NodeGeom mygm = {0.0, 0.0, 0.0, 0.0};
      if (has) { mylk = a.link[mynode]; mygm = a.geom[mynode]; }
      if (has) {
        if (mylk.nchild != 0u) {
          const double* mm = a.mom64 + (size_t)mynode * 16;
          smom[lane][0] = mm[3];
#pragma unroll
          for (int j = 0; j < 3; ++j) smom[lane][1 + j] = mm[5 + j];
#pragma unroll
          for (int j = 0; j < 6; ++j) smom[lane][4 + j] = mm[8 + j];
        } else {
          smom[lane][0] = a.q[mylk.first];
        }
      }
      __syncthreads();
      sp -= m;
      for (uint32_t k = 0; k < m && !overflow; ++k) {
        const uint32_t node = lane_bcast(mynode, (int)k);
        const uint32_t first = lane_bcast(mylk.first, (int)k);
        const uint32_t nchild = lane_bcast(mylk.nchild, (int)k);
        const double cx = lane_bcast(mygm.cx, (int)k), cy = lane_bcast(mygm.cy, (int)k),
                     cz = lane_bcast(mygm.cz, (int)k), b2 = lane_bcast(mygm.b2, (int)k);
        if (nchild != 0u) {
          bool fail;
          {
#pragma clang fp contract(off)
            const double dx = xt - cx, dy = yt - cy, dz = zt - cz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            fail = valid && !(a.theta2 * d2 > b2);
          }
          if (__ballot(fail) != 0ull) {
            if (sp + nchild > (uint32_t)(kStackCap + kStackSlack)) { overflow = true; break; }
            if ((uint32_t)lane < nchild) stack[sp + lane] = first + (uint32_t)lane;
            sp += nchild;
            continue;
          }
        }
        double d[3] = {xt - cx, yt - cy, zt - cz};
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps2;
        double D[3], gq, gp[3], M[6];
        if (nchild == 0u) {   // leaf: P2P with source particle `first`
          const double q = smom[k][0];
          const bool self = first == tid || !valid;
          adjoint_term_p2p(d, self ? 1.0 : r2, gphi, u, q, D, gq);
          if (self) { D[0] = D[1] = D[2] = 0.0; gq = 0.0; }
          tD[0] += D[0]; tD[1] += D[1]; tD[2] += D[2];
          // values -D (to the source particle's position) and gq; lane 16 r
          // adds value row_value(r)
          const double sv = wsum4(-D[0], -D[1], -D[2], gq);
          if ((lane & 15) == 0) {
            const int j = row_value(lane >> 4);
            double* dst = j == 0 ? a.gx : j == 1 ? a.gy : j == 2 ? a.gz : a.gq;
            unsafeAtomicAdd(&dst[first], sv);
          }
          continue;
        }
        const double q = smom[k][0];
        const double p[3] = {smom[k][1], smom[k][2], smom[k][3]};
        const double qt[6] = {smom[k][4], smom[k][5], smom[k][6], smom[k][7], smom[k][8], smom[k][9]};
        adjoint_term(d, r2, gphi, u, q, p, qt, D, gq, gp, M);
        if (!valid) {
          D[0] = D[1] = D[2] = 0.0; gq = 0.0; gp[0] = gp[1] = gp[2] = 0.0;
#pragma unroll
          for (int j = 0; j < 6; ++j) M[j] = 0.0;
        }
        tD[0] += D[0]; tD[1] += D[1]; tD[2] += D[2];
        // the node's 13 sums (slots 0..12): register c holds values 4c..4c+3,
        // and lane 16 r + c adds value 4c + row_value(r)
        const double s0 = wsum4(gq, gp[0], gp[1], gp[2]);
        const double s1 = wsum4(M[0], M[1], M[2], M[3]);
        const double s2 = wsum4(M[4], M[5], D[0], D[1]);
        const double s3 = wsum4(D[2], 0.0, 0.0, 0.0);
        const int c = lane & 15;
        const int j = 4 * c + row_value(lane >> 4);
        if (c < 4 && j < 13) {
          const double sv = c == 0 ? s0 : c == 1 ? s1 : c == 2 ? s2 : s3;
          unsafeAtomicAdd(&a.nadj[(size_t)node * 16 + j], sv);
        }
      }
      __syncthreads();
    }
    if (overflow && lane == 0) atomicOr(&a.sc->flags, kFlagStackOverflow);
    if (valid) {
      unsafeAtomicAdd(&a.gx[tid], tD[0]);
      unsafeAtomicAdd(&a.gy[tid], tD[1]);
      unsafeAtomicAdd(&a.gz[tid], tD[2]);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------- frozen lists
// Count (kRecord = false) or write (true) each group's interaction list: the
// node of every leaf and accepted twig in traversal order.  Both passes make
// the same decisions, so entries land in [off[g], off[g] + count[g]).
template <bool kRecord>
__global__ __launch_bounds__(kWave) void list_walk_kernel(ListArgs a) {
  __shared__ uint32_t stack[kStackCap + kStackSlack];
  __shared__ uint32_t gsh;
  const int lane = threadIdx.x;
  const uint32_t ngroups = a.sc->ngroups;
  for (;;) {
    if (lane == 0) gsh = atomicAdd(a.queue, 1u);
    __syncthreads();
    const uint32_t g = gsh;
    __syncthreads();
    if (g >= ngroups) break;
    const uint32_t a0 = a.gfirst[g], cnt = a.gcount[g];
    const bool valid = (uint32_t)lane < cnt;
    const uint32_t tid = a0 + (valid ? (uint32_t)lane : 0u);
    const double xt = a.x[tid], yt = a.y[tid], zt = a.z[tid];
    const uint32_t root = a.sys ? a.sysroot[a.sys[a0]] : 0u;
    uint32_t k = 0;
    const uint32_t base = kRecord ? a.off[g] : 0u;
    auto entry = [&](uint32_t node, uint32_t, uint32_t, double, double, double) {
      if (kRecord && lane == 0) a.entries[base + k] = node;
      ++k;
    };
    const bool overflow = traverse_group(root, xt, yt, zt, valid, a.theta2, a.link, a.geom, stack,
                                         entry);
    if (overflow && lane == 0) atomicOr(&a.sc->flags, kFlagStackOverflow);
    if (!kRecord && lane == 0) a.count[g] = k;
    __syncthreads();
  }
}

// 64-bit total of the list lengths (the 32-bit offsets must not wrap).
__global__ void list_total_kernel(const uint32_t* __restrict__ count,
                                  unsigned long long* __restrict__ total,
                                  const DevScalars* __restrict__ sc) {
  const int64_t ng = sc->ngroups;
  unsigned long long v = 0;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng;
       g += (int64_t)gridDim.x * blockDim.x)
    v += count[g];
  v = wave_sum(v);
  if (lane_id() == 0 && v) atomicAdd(total, v);
}

__global__ void list_clear_kernel(unsigned long long* total, uint32_t* queue) {
  if (threadIdx.x == 0) { *total = 0ull; *queue = 0u; }
}

__global__ void queue_clear_kernel(uint32_t* queue) {
  if (threadIdx.x == 0) *queue = 0u;
}

// Leaf geometry from (new) particle data: centre = particle, last slot = charge.
__global__ void leaf_geom_kernel(const NodeLink* __restrict__ link, const double* __restrict__ x,
                                 const double* __restrict__ y, const double* __restrict__ z,
                                 const double* __restrict__ q, NodeGeom* __restrict__ geom,
                                 const DevScalars* __restrict__ sc) {
  if (broken(sc)) return;
  const int64_t m = sc->nnodes;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const NodeLink lk = link[i];
    if (lk.nchild != 0u) continue;
    NodeGeom gm;
    gm.cx = x[lk.first]; gm.cy = y[lk.first]; gm.cz = z[lk.first]; gm.b2 = q[lk.first];
    geom[i] = gm;
  }
}

// Fields of the frozen lists, FP64, one group per wavefront, lane = target.
// M2P from the reference-layout moments: phi = q/r - p.d/r^3 - tr/r^3 +
// 3 d.Qt.d/r^5 and E = -grad phi (module_coulomb_kernels.f90:60-145, with
// r^2 = d^2 + eps^2); P2P phi = q/r, E = q d/r^3, self excluded by index.
__global__ __launch_bounds__(kWave) void replay_kernel(ReplayArgs a) {
  __shared__ uint32_t gsh;
  const int lane = threadIdx.x;
  const uint32_t ngroups = a.sc->ngroups;
  for (;;) {
    if (lane == 0) gsh = atomicAdd(a.queue, 1u);
    __syncthreads();
    const uint32_t g = gsh;
    __syncthreads();
    if (g >= ngroups) break;
    const uint32_t a0 = a.gfirst[g], cnt = a.gcount[g];
    const bool valid = (uint32_t)lane < cnt;
    const uint32_t tid = a0 + (valid ? (uint32_t)lane : 0u);
    const double xt = a.x[tid], yt = a.y[tid], zt = a.z[tid];
    double e[3] = {0.0, 0.0, 0.0}, ph = 0.0;
    const uint32_t b = a.off[g], nent = a.count[g];
    for (uint32_t k = 0; k < nent; ++k) {
      const uint32_t node = a.entries[b + k];
      const NodeLink lk = a.link[node];
      const NodeGeom gm = a.geom[node];
      const double d[3] = {xt - gm.cx, yt - gm.cy, zt - gm.cz};
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + a.eps2;
      if (lk.nchild == 0u) {
        if (lk.first == tid || !valid) continue;
        const double ri = 1.0 / sqrt(r2);
        const double qr3 = gm.b2 * ri * ri * ri;
        ph += gm.b2 * ri;
        for (int c = 0; c < 3; ++c) e[c] += qr3 * d[c];
        continue;
      }
      const double* mm = a.mom64 + (size_t)node * 16;
      const double q = mm[3];
      const double p[3] = {mm[5], mm[6], mm[7]};
      const double hxy = 0.5 * mm[11], hyz = 0.5 * mm[12], hzx = 0.5 * mm[13];
      const double Qd[3] = {mm[8] * d[0] + hxy * d[1] + hzx * d[2],
                            hxy * d[0] + mm[9] * d[1] + hyz * d[2],
                            hzx * d[0] + hyz * d[1] + mm[10] * d[2]};
      const double tr = mm[8] + mm[9] + mm[10];
      const double ri = 1.0 / sqrt(r2);
      const double r3 = ri * ri * ri, r5 = r3 * ri * ri, r7 = r5 * ri * ri;
      const double pd = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
      const double dQd = d[0] * Qd[0] + d[1] * Qd[1] + d[2] * Qd[2];
      ph += q * ri - pd * r3 - tr * r3 + 3.0 * dQd * r5;
      for (int c = 0; c < 3; ++c)
        e[c] += q * d[c] * r3 + p[c] * r3 - 3.0 * pd * d[c] * r5 - 3.0 * tr * d[c] * r5 -
                6.0 * Qd[c] * r5 + 15.0 * dQd * d[c] * r7;
    }
    if (valid) {
      a.ex[tid] = a.sign * e[0];
      a.ey[tid] = a.sign * e[1];
      a.ez[tid] = a.sign * e[2];
      a.pot[tid] = a.sign * ph;
    }
    __syncthreads();
  }
}

// Centre adjoint of every twig: dL/dc = q gp + M p - sum D (slots 13..15),
// and parent links for the particle pass.
__global__ void adjoint_nodes_kernel(const NodeLink* __restrict__ link, const double* __restrict__ mom64,
                                     double* __restrict__ nadj, uint32_t* __restrict__ parent,
                                     uint32_t* __restrict__ leafof, const DevScalars* __restrict__ sc) {
  if (broken(sc)) return;
  const int64_t m = sc->nnodes;
  const int64_t nroot = sc->level_offset[1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const NodeLink lk = link[i];
    if (i < nroot) parent[i] = 0xFFFFFFFFu;
    if (lk.nchild == 0u) { leafof[lk.first] = (uint32_t)i; continue; }
    for (uint32_t c = 0; c < lk.nchild; ++c) parent[lk.first + c] = (uint32_t)i;
    double* o = nadj + (size_t)i * 16;
    const double* mm = mom64 + (size_t)i * 16;
    const double Mxx = o[4], Myy = o[5], Mzz = o[6], Mxy = o[7], Myz = o[8], Mzx = o[9];
    const double px = mm[5], py = mm[6], pz = mm[7], q = mm[3];
    o[13] = q * o[1] + (Mxx * px + Mxy * py + Mzx * pz) - o[10];
    o[14] = q * o[2] + (Mxy * px + Myy * py + Myz * pz) - o[11];
    o[15] = q * o[3] + (Mzx * px + Myz * py + Mzz * pz) - o[12];
  }
}

// Each particle collects the adjoints of its twig ancestors.
__global__ void adjoint_particles_kernel(int64_t n, const double* __restrict__ x,
                                         const double* __restrict__ y, const double* __restrict__ z,
                                         const double* __restrict__ q,
                                         const double* __restrict__ mom64,
                                         const double* __restrict__ nadj,
                                         const uint32_t* __restrict__ parent,
                                         const uint32_t* __restrict__ leafof,
                                         double* __restrict__ gx, double* __restrict__ gy,
                                         double* __restrict__ gz, double* __restrict__ gq,
                                         const DevScalars* __restrict__ sc) {
  if (broken(sc)) return;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const double xj[3] = {x[j], y[j], z[j]};
    const double qj = q[j];
    const double sg = qj > 0.0 ? 1.0 : (qj < 0.0 ? -1.0 : 0.0);
    double dq = 0.0, dx[3] = {0.0, 0.0, 0.0};
    uint32_t node = parent[leafof[j]];
    for (int guard = 0; node != 0xFFFFFFFFu && guard <= kMaxLevel; ++guard) {
      const double* o = nadj + (size_t)node * 16;
      const double* mm = mom64 + (size_t)node * 16;
      const double s[3] = {mm[0] - xj[0], mm[1] - xj[1], mm[2] - xj[2]};
      const double Ms[3] = {o[4] * s[0] + o[7] * s[1] + o[9] * s[2],
                            o[7] * s[0] + o[5] * s[1] + o[8] * s[2],
                            o[9] * s[0] + o[8] * s[1] + o[6] * s[2]};
      dq += o[0] + o[1] * s[0] + o[2] * s[1] + o[3] * s[2] +
            0.5 * (s[0] * Ms[0] + s[1] * Ms[1] + s[2] * Ms[2]);
#pragma unroll
      for (int k = 0; k < 3; ++k) dx[k] += -qj * (o[1 + k] + Ms[k]);
      const double A = mm[4];
      if (A != 0.0) {
        dq += sg / A * -(s[0] * o[13] + s[1] * o[14] + s[2] * o[15]);
#pragma unroll
        for (int k = 0; k < 3; ++k) dx[k] += fabs(qj) / A * o[13 + k];
      }
      node = parent[node];
    }
    gq[j] += dq;
    gx[j] += dx[0];
    gy[j] += dx[1];
    gz[j] += dx[2];
  }
}

unsigned grid_for(int64_t n) {
  int64_t b = (n + kT - 1) / kT;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (unsigned)b;
}

}  // namespace

void launch_lists(hipStream_t s, const ListArgs& a, bool record, int blocks,
                  unsigned long long* total) {
  if (record) {
    queue_clear_kernel<<<1, 64, 0, s>>>(a.queue);
    list_walk_kernel<true><<<blocks, kWave, 0, s>>>(a);
  } else {
    list_clear_kernel<<<1, 64, 0, s>>>(total, a.queue);
    list_walk_kernel<false><<<blocks, kWave, 0, s>>>(a);
    list_total_kernel<<<256, kT, 0, s>>>(a.count, total, a.sc);
  }
}

void launch_leaf_geom(hipStream_t s, const NodeLink* link, const double* x, const double* y,
                      const double* z, const double* q, NodeGeom* geom, int64_t cap,
                      const DevScalars* sc) {
  leaf_geom_kernel<<<grid_for(cap), kT, 0, s>>>(link, x, y, z, q, geom, sc);
}

void launch_replay(hipStream_t s, const ReplayArgs& a, int blocks) {
  queue_clear_kernel<<<1, 64, 0, s>>>(a.queue);
  replay_kernel<<<blocks, kWave, 0, s>>>(a);
}

void launch_adjoint(hipStream_t s, const AdjointArgs& p, int blocks) {
  const int64_t n = p.n;
  adjoint_clear_kernel<<<grid_for(p.cap * 16 > n ? p.cap * 16 : n), kT, 0, s>>>(
      n, p.cap, p.gx, p.gy, p.gz, p.gq, p.nadj, p.queue, p.sc);
  AdjArgs a;
  a.x = p.x; a.y = p.y; a.z = p.z; a.q = p.q;
  a.geom = p.geom; a.link = p.link; a.mom64 = p.mom64;
  a.gfirst = p.gfirst; a.gcount = p.gcount; a.sys = p.sys; a.sysroot = p.sysroot;
  a.sc = p.sc; a.theta2 = p.theta2; a.eps2 = p.eps2; a.sign = p.sign;
  a.gex = p.gex; a.gey = p.gey; a.gez = p.gez; a.gpot = p.gpot;
  a.gx = p.gx; a.gy = p.gy; a.gz = p.gz; a.gq = p.gq;
  a.nadj = p.nadj; a.queue = p.queue;
  adjoint_walk_kernel<<<blocks, kWave, 0, s>>>(a);
  adjoint_nodes_kernel<<<grid_for(p.cap), kT, 0, s>>>(p.link, p.mom64, p.nadj, p.parent, p.leafof,
                                                      p.sc);
  adjoint_particles_kernel<<<grid_for(n), kT, 0, s>>>(n, p.x, p.y, p.z, p.q, p.mom64, p.nadj,
                                                      p.parent, p.leafof, p.gx, p.gy, p.gz, p.gq,
                                                      p.sc);
}

}  // namespace pepc
