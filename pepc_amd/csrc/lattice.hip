// lattice.hip -- periodic boundaries: lattice coefficients, multipole moments
// of the central box, and the far-field (lattice) contribution per particle.
//
// Scheme of the reference's fast-multipole framework
// (coulomb/module_fmm_framework.f90; Challacombe, White and Head-Gordon,
// J. Chem. Phys. 107, 10131 (1997); Kudin and Scuseria, Chem. Phys. Lett.
// 283, 61 (1998)).  The near field (the mirror boxes within one shell) is the
// tree walk over every lattice vector (walk.hip); everything beyond it enters
// through
//   ML        lattice coefficients (Taylor table, l <= 40), fixed per lattice;
//   omega     multipole moments (l <= 20) of the central box about its centre,
//             particles plus the fictitious charges that cancel the box
//             dipole, monopole removed;
//   mu        M2L(ML, omega), the lattice's local expansion at the centre;
// and per particle the value and gradient of mu shifted to the particle
// (L2L to order 1) plus the fictitious charges of every near box.
//
// Tables hold l = 0..lmax, m = 0..l at l(l+1)/2 + m (tblinv, :863-873);
// negative orders are (-1)^m conj (tbl, :838-858).  The harmonics are the
// reference's: O_lm = r^l P_l^m(cos t) / (l+m)! e^{-i m p} and
// M_lm = r^-(l+1) P_l^m(cos t) (l-m)! e^{i m p}, P without the Condon-Shortley
// phase (:659-734, :1196-1245).
//
// Work per launch: the coefficient kernels run once per lattice; the moments
// kernel is N x 231 complex terms (one grid row per order m, recurrences in
// Cartesian form instead of Legendre evaluations per term); the field kernel
// is N x 231 terms against mu in LDS plus the neighbour boxes' fictitious
// charges merged by lattice point (46 points for one mirror layer instead of
// 4 x 27 pairs: where all four charges meet they cancel).
#include "kernels.h"

namespace pepc {

namespace {

constexpr int kLM = kLatticeLM;     // Lmax_multipole, :70
constexpr int kLT = 2 * kLM;        // Lmax_taylor, :71
constexpr int kNM = kLatticeNM;     // 231
constexpr int kNT = kLatticeNT;     // 861

__host__ __device__ constexpr int tix(int l, int m) { return l * (l + 1) / 2 + m; }

__device__ __forceinline__ void tdecode(int t, int& l, int& m) {
  int ll = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (tix(ll + 1, 0) <= t) ++ll;
  while (tix(ll, 0) > t) --ll;
  l = ll;
  m = t - tix(ll, 0);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// tbl(A, l, m, lmax) for |m| <= l: stored entry, or (-1)^m conj for m < 0;
// zero beyond lmax or |m| > l.
__device__ __forceinline__ double2 tbl(const double2* a, int l, int m, int lmax) {
  if (l > lmax || m > l || -m > l) return make_double2(0.0, 0.0);
  if (m >= 0) return a[tix(l, m)];
  const double2 v = a[tix(l, -m)];
  return (m & 1) ? make_double2(-v.x, v.y) : make_double2(v.x, -v.y);
}

// LegendreP without the Condon-Shortley phase (:1196-1245).
__device__ double legendre_p(int l, int m, double x) {
  double pmm = 1.0, v;
  if (m > 0) {
    const double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; ++i) {
      pmm = -pmm * fact * somx2;
      fact += 2.0;
    }
  }
  if (l == m) {
    v = pmm;
  } else {
    double pmmp1 = x * (2.0 * m + 1.0) * pmm;
    if (l == m + 1) {
      v = pmmp1;
    } else {
      double pll = 0.0;
      for (int ll = m + 2; ll <= l; ++ll) {
        pll = (x * (2.0 * ll - 1.0) * pmmp1 - (ll + m - 1.0) * pmm) / (double)(ll - m);
        pmm = pmmp1;
        pmmp1 = pll;
      }
      v = pll;
    }
  }
  return (m & 1) ? -v : v;
}

struct Sph { double r, ct, ph; };

__device__ __forceinline__ Sph to_sph(double x, double y, double z) {   // :1083-1103
  Sph s;
  s.r = sqrt(x * x + y * y + z * z);
  s.ct = s.r == 0.0 ? 1.0 : z / s.r;
  s.ph = (x == 0.0 && y == 0.0) ? 0.0 : atan2(y, x);
  return s;
}

// OMultipole (:695-715) for 0 <= m <= l.
__device__ double2 o_multipole(int l, int m, Sph s, const double* fact) {
  const double rl = (s.r > 0.0 || l > 0) ? pow(s.r, (double)l) : 1.0;
  const double a = rl * (legendre_p(l, m, s.ct) / fact[l + m]);
  return make_double2(a * cos((double)m * s.ph), -a * sin((double)m * s.ph));
}

// MTaylor (:720-734) for 0 <= m <= l.
__device__ double2 m_taylor(int l, int m, Sph s, const double* fact) {
  const double a = 1.0 / pow(s.r, (double)(l + 1)) * (legendre_p(l, m, s.ct) * fact[l - m]);
  return make_double2(a * cos((double)m * s.ph), a * sin((double)m * s.ph));
}

// ---------------------------------------------------------------------------
// Lattice coefficients (calc_lattice_coefficients, :154-213).
// One block per entry of M* (kNM blocks, MstarFunc :1005-1036) and of L*
// (kNT blocks, LstarFunc :1042-1077): one term per thread, terms sorted by
// magnitude in LDS (real and imaginary parts separately), then summed from
// the smallest up, as the reference does to avoid cancellation.
constexpr int kTermCap = 1024;

__global__ __launch_bounds__(kTermCap) void lattice_terms_kernel(const double* __restrict__ vec,
                                                                 int nb, double scale,
                                                                 const double* __restrict__ fact,
                                                                 double2* __restrict__ mstar,
                                                                 double2* __restrict__ lstar) {
  __shared__ double re[kTermCap], im[kTermCap];
  const bool is_m = blockIdx.x < (unsigned)kNM;
  const int t = is_m ? (int)blockIdx.x : (int)blockIdx.x - kNM;
  int l, m;
  tdecode(t, l, m);
  const int nterm = is_m ? nb : (nb - 1) * nb;
  const int i = threadIdx.x;
  double2 v = make_double2(0.0, 0.0);
  if (i < nterm) {
    if (is_m) {
      v = o_multipole(l, m, to_sph(-vec[3 * i], -vec[3 * i + 1], -vec[3 * i + 2]), fact);
    } else {
      const int a = i / nb, b = i - a * nb;   // a: outer box (central box excluded)
      const double cx = scale * vec[3 * a] + vec[3 * b];
      const double cy = scale * vec[3 * a + 1] + vec[3 * b + 1];
      const double cz = scale * vec[3 * a + 2] + vec[3 * b + 2];
      v = m_taylor(l, m, to_sph(cx, cy, cz), fact);
    }
  }
  re[i] = v.x;
  im[i] = v.y;
  __syncthreads();
  // bitonic sort of both arrays by magnitude (zeros of the padding first)
  for (int k = 2; k <= kTermCap; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = i ^ j;
      if (p > i) {
        const bool up = (i & k) == 0;
        double a = re[i], b = re[p];
        if ((fabs(a) > fabs(b)) == up) { re[i] = b; re[p] = a; }
        a = im[i]; b = im[p];
        if ((fabs(a) > fabs(b)) == up) { im[i] = b; im[p] = a; }
      }
      __syncthreads();
    }
  }
  if (i == 0) {
    double rp = 0.0, ip = 0.0;
    for (int k = 0; k < kTermCap; ++k) { rp += re[k]; ip += im[k]; }
    if (is_m) mstar[t] = make_double2(rp, ip);
    else lstar[t] = make_double2(rp, ip);
  }
}

// More terms than one block sorts in LDS (mirror_box_layers >= 2: 124 x 125
// or 342 x 343 L* terms per entry): every term is added in double-double
// arithmetic (TwoSum, Knuth; contraction is off in this build), per thread
// over a strided run, then in a fixed tree over the block, and rounded once.
// The result is the exactly rounded sum up to about 2^-100 relative, which
// is at least as accurate as the reference's smallest-first order.
struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, double b) {
  const double s = a.hi + b;
  const double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b - bb);
  e += a.lo;
  const double h = s + e;
  return DD{h, e - (h - s)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  const double h = s + e;
  return DD{h, e - (h - s)};
}

constexpr int kTermDD = 1024;

__global__ __launch_bounds__(kTermDD) void lattice_terms_dd_kernel(const double* __restrict__ vec,
                                                                   int nb, double scale,
                                                                   const double* __restrict__ fact,
                                                                   double2* __restrict__ mstar,
                                                                   double2* __restrict__ lstar) {
  __shared__ DD sre[kTermDD], sim[kTermDD];
  const bool is_m = blockIdx.x < (unsigned)kNM;
  const int t = is_m ? (int)blockIdx.x : (int)blockIdx.x - kNM;
  int l, m;
  tdecode(t, l, m);
  const int64_t nterm = is_m ? (int64_t)nb : (int64_t)(nb - 1) * nb;
  DD re{0.0, 0.0}, im{0.0, 0.0};
  for (int64_t i = threadIdx.x; i < nterm; i += kTermDD) {
    double2 v;
    if (is_m) {
      v = o_multipole(l, m, to_sph(-vec[3 * i], -vec[3 * i + 1], -vec[3 * i + 2]), fact);
    } else {
      const int64_t a = i / nb, b = i - a * nb;   // a: outer box (central box excluded)
      const double cx = scale * vec[3 * a] + vec[3 * b];
      const double cy = scale * vec[3 * a + 1] + vec[3 * b + 1];
      const double cz = scale * vec[3 * a + 2] + vec[3 * b + 2];
      v = m_taylor(l, m, to_sph(cx, cy, cz), fact);
    }
    re = dd_add(re, v.x);
    im = dd_add(im, v.y);
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int h = kTermDD / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sre[threadIdx.x] = dd_add(sre[threadIdx.x], sre[threadIdx.x + h]);
      sim[threadIdx.x] = dd_add(sim[threadIdx.x], sim[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double2 r = make_double2(sre[0].hi + sre[0].lo, sim[0].hi + sim[0].lo);
    if (is_m) mstar[t] = r;
    else lstar[t] = r;
  }
}

// ML = L*; then maxiter times ML = M2L(UL(ML), M*) + L* (UL divides order l by
// scale^(l+1), :982-1000; M2L :912-938).  One block, one entry per thread.
__global__ __launch_bounds__(1024) void lattice_iterate_kernel(const double2* __restrict__ mstar,
                                                              const double2* __restrict__ lstar,
                                                              double scale, int maxiter,
                                                              double2* __restrict__ ml) {
  __shared__ double2 ul[kNT];
  __shared__ double2 ms[kNM];
  const int t = threadIdx.x;
  int l = 0, m = 0;
  if (t < kNT) tdecode(t, l, m);
  if (t < kNM) ms[t] = mstar[t];
  const double2 ls = t < kNT ? lstar[t] : make_double2(0.0, 0.0);
  const double div = pow(scale, (double)(l + 1));
  double2 cur = ls;
  for (int it = 0; it < maxiter; ++it) {
    if (t < kNT) ul[t] = make_double2(cur.x / div, cur.y / div);
    __syncthreads();
    if (t < kNT) {
      double2 acc = make_double2(0.0, 0.0);
      const int jmax = kLM < kLT - l ? kLM : kLT - l;
      for (int j = 0; j <= jmax; ++j) {
        const double sj = (j & 1) ? -1.0 : 1.0;
        for (int k = -j; k <= j; ++k)
          acc = cadd(acc, cmul(cscale(sj, tbl(ul, j + l, k + m, kLT)), tbl(ms, j, k, kLM)));
      }
      cur = cadd(acc, ls);
    }
    __syncthreads();
  }
  if (t < kNT) ml[t] = cur;
}

// ---------------------------------------------------------------------------
// Per-step moments of the central box about its centre (calc_box_dipole
// :605-631, calc_omega_tilde :514-600).  Grid (bx, kLM + 1): row M (a
// template parameter, so every loop below has a fixed trip count and the
// recurrence constants fold) accumulates q O_lM for l = M..kLM with
//   O_MM = (x - i y)^M / (2^M M!),
//   O_lM = ((2l-1) z O_{l-1,M} - r^2 O_{l-2,M}) / ((l-M)(l+M)),
// which equal the reference's r^l P_l^M / (l+M)! e^{-i M p}.  Row 0 also
// sums the dipole q r and the quadrupole trace q r.r (calc_box_dipole).  Every thread handles a strided run of particles (the
// next one's coordinates loaded while the current one is evaluated); each
// block writes its partial sums, which lattice_omega_kernel adds in block
// order (deterministic).
constexpr int kMomT = 256;
constexpr int kMomRed = 2 * (kLM + 1) + 4;

template <int M>
__device__ __forceinline__ void moments_row(int64_t n, const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const double* __restrict__ z,
                                            const double* __restrict__ q, const LatticeGeom& g,
                                            double2* __restrict__ part, double* __restrict__ dpart,
                                            double (*red)[kMomRed]) {
#pragma clang fp contract(fast)   // tolerance-checked sums: fused multiply-adds allowed
  constexpr int K = kLM + 1 - M;
  double ar[K], ai[K], dp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < K; ++j) { ar[j] = 0.0; ai[j] = 0.0; }
  const int64_t stride = (int64_t)gridDim.x * kMomT;
  int64_t i = blockIdx.x * (int64_t)kMomT + threadIdx.x;
  double nx = 0.0, ny = 0.0, nz = 0.0, nq = 0.0;
  if (i < n) { nx = x[i]; ny = y[i]; nz = z[i]; nq = q[i]; }
  while (i < n) {
    const double rx = nx - g.center[0], ry = ny - g.center[1], rz = nz - g.center[2];
    const double qi = nq;
    const int64_t i2 = i + stride;
    if (i2 < n) { nx = x[i2]; ny = y[i2]; nz = z[i2]; nq = q[i2]; }
    i = i2;
    const double r2 = rx * rx + ry * ry + rz * rz;
    if constexpr (M == 0) {
      dp[0] += qi * rx; dp[1] += qi * ry; dp[2] += qi * rz; dp[3] += qi * r2;
    }
    double wr = 1.0, wi = 0.0;
#pragma unroll
    for (int k = 1; k <= M; ++k) {
      const double h = 0.5 / (double)k;
      const double tr = (wr * rx + wi * ry) * h, ti = (wi * rx - wr * ry) * h;
      wr = tr; wi = ti;
    }
    ar[0] += qi * wr;
    ai[0] += qi * wi;
    double p1r = wr, p1i = wi, p2r = 0.0, p2i = 0.0;
#pragma unroll
    for (int j = 1; j < K; ++j) {
      const int l = M + j;
      const double a = (double)(2 * l - 1) * rz;
      const double inv = 1.0 / (double)((l - M) * (l + M));
      const double cr = (a * p1r - r2 * p2r) * inv, ci = (a * p1i - r2 * p2i) * inv;
      ar[j] += qi * cr;
      ai[j] += qi * ci;
      p2r = p1r; p2i = p1i; p1r = cr; p1i = ci;
    }
  }
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const double sr = wave_sum(ar[j]), si = wave_sum(ai[j]);
    if (lane_id() == 0) { red[w][2 * j] = sr; red[w][2 * j + 1] = si; }
  }
  if constexpr (M == 0) {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const double sd = wave_sum(dp[d]);
      if (lane_id() == 0) red[w][2 * K + d] = sd;
    }
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k < 2 * K + (M == 0 ? 4 : 0)) {
    double sum = 0.0;
    for (int v = 0; v < kMomT / kWave; ++v) sum += red[v][k];
    if (k < 2 * K) {
      double* dst = reinterpret_cast<double*>(&part[(size_t)blockIdx.x * kNM + tix(M + (k >> 1), M)]);
      dst[k & 1] = sum;
    } else {
      dpart[(size_t)blockIdx.x * 4 + (k - 2 * K)] = sum;
    }
  }
}

template <int M>
__device__ __forceinline__ void moments_dispatch(int row, int64_t n, const double* x, const double* y,
                                                 const double* z, const double* q,
                                                 const LatticeGeom& g, double2* part, double* dpart,
                                                 double (*red)[kMomRed]) {
  if (row == M) moments_row<M>(n, x, y, z, q, g, part, dpart, red);
  else if constexpr (M < kLM) moments_dispatch<M + 1>(row, n, x, y, z, q, g, part, dpart, red);
}

__global__ __launch_bounds__(kMomT) void lattice_moments_kernel(int64_t n, const double* __restrict__ x,
                                                                const double* __restrict__ y,
                                                                const double* __restrict__ z,
                                                                const double* __restrict__ q,
                                                                LatticeGeom g,
                                                                double2* __restrict__ part,
                                                                double* __restrict__ dpart) {
  __shared__ double red[kMomT / kWave][kMomRed];
  moments_dispatch<0>((int)blockIdx.y, n, x, y, z, q, g, part, dpart, red);
}

// omega (and the dipole and quadrupole trace) from the per-block partials:
// one wave per entry, lane j adds blocks j, j+64, ... in order, then a fixed
// butterfly.
__global__ __launch_bounds__(kWave) void lattice_omega_kernel(const double2* __restrict__ part,
                                                             const double* __restrict__ dpart,
                                                             int nblk, double2* __restrict__ om,
                                                             double* __restrict__ dip) {
  const int e = blockIdx.x;
  const int lane = threadIdx.x;
  double sr = 0.0, si = 0.0;
  if (e < kNM) {
    for (int b = lane; b < nblk; b += kWave) {
      const double2 v = part[(size_t)b * kNM + e];
      sr += v.x; si += v.y;
    }
  } else {
    for (int b = lane; b < nblk; b += kWave) sr += dpart[(size_t)b * 4 + (e - kNM)];
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if (lane == 0) {
    if (e < kNM) om[e] = make_double2(sr, si);
    else dip[e - kNM] = sr;
  }
}

// With the fictitious-charge correction (Kudin, :535-554) adds, for each
// periodic axis p, a charge -D_p / |t_p| at origin + t_p, and their negated
// sum at the origin, to omega; removes the monopole (:570-573) and forms
// mu = M2L(ML, omega) for l <= kLM (the orders the field needs), four
// threads per entry (j = part mod 4), their sums added in order.  One block.
constexpr int kMuT = 1024;

__global__ __launch_bounds__(kMuT) void lattice_mu_kernel(const double2* __restrict__ om_g,
                                                         const double* __restrict__ dip_g,
                                                         LatticeGeom g, const double2* __restrict__ ml_g,
                                                         const double* __restrict__ fact,
                                                         double2* __restrict__ mu,
                                                         double* __restrict__ fict,
                                                         double* __restrict__ coef) {
  __shared__ double2 om[kNM];
  __shared__ double2 mus[kNM];
  __shared__ double2 ml[kNT];
  __shared__ double2 acc4[kNM][4];
  __shared__ double fq[4], fpos[4][3];
  __shared__ int nf;
  const int t = threadIdx.x;
  for (int k = t; k < kNT; k += kMuT) ml[k] = ml_g[k];
  if (t == 0) {
    int k = 0;
    if (g.correction == kExtrinsicFictCharge) {
      double sum = 0.0;
      for (int p = 0; p < 3; ++p) {
        if (!g.per[p]) continue;
        const double* tv = g.lat + 3 * p;
        for (int d = 0; d < 3; ++d) fpos[k][d] = g.origin[d] + tv[d];
        fq[k] = -dip_g[p] / sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
        sum += fq[k];
        ++k;
      }
      for (int d = 0; d < 3; ++d) fpos[k][d] = g.origin[d];
      fq[k] = -sum;
      ++k;
    }
    nf = k;
  }
  __syncthreads();
  // the neighbour boxes' fictitious charges merged by lattice point: charge
  // f sits at origin + t_p (axis p) or at the origin, so every contribution
  // to a point is one of this step's charges
  for (int k = t; k < g.nfpt; k += kMuT) {
    const int* e = g.fpt + 4 * k;
    double qk = 0.0;
    for (int f = 0; f < nf; ++f)
      if ((e[3] >> f) & 1) qk += fq[f];
    for (int d = 0; d < 3; ++d)
      g.fpos[4 * k + d] = g.origin[d] + (double)e[0] * g.lat[d] + (double)e[1] * g.lat[3 + d] +
                          (double)e[2] * g.lat[6 + d];
    g.fpos[4 * k + 3] = qk;
  }
  if (t < kNM) {
    int l, m;
    tdecode(t, l, m);
    double2 v = om_g[t];
    for (int f = 0; f < nf; ++f) {
      const Sph s = to_sph(fpos[f][0] - g.center[0], fpos[f][1] - g.center[1], fpos[f][2] - g.center[2]);
      v = cadd(v, cscale(fq[f], o_multipole(l, m, s, fact)));
    }
    if (t == 0) v = make_double2(0.0, 0.0);
    om[t] = v;
  }
  __syncthreads();
  const int e = t >> 2, part = t & 3;
  if (e < kNM) {
    int l, m;
    tdecode(e, l, m);
    double2 acc = make_double2(0.0, 0.0);
    for (int j = part; j <= kLM; j += 4) {
      const double sj = (j & 1) ? -1.0 : 1.0;
      for (int k = -j; k <= j; ++k)
        acc = cadd(acc, cmul(cscale(sj, tbl(ml, j + l, k + m, kLT)), tbl(om, j, k, kLM)));
    }
    acc4[e][part] = acc;
  }
  __syncthreads();
  if (t < kNM) {
    const double2 v = cadd(cadd(cadd(acc4[t][0], acc4[t][1]), acc4[t][2]), acc4[t][3]);
    mu[t] = v;
    mus[t] = v;
  }
  if (t < 4 * 4) fict[t] = (t / 4) < nf ? (t % 4 < 3 ? fpos[t / 4][t % 4] : fq[t / 4]) : 0.0;
  if (t == 0) fict[16] = (double)nf;
  if (t >= 17 && t < kLatticeFict) fict[t] = dip_g[t - 17];   // box dipole, quadrupole trace
  __syncthreads();
  // Field coefficients per entry (l, m): the L00, L10 and L11 sums of the
  // field kernel are linear in O_lm = (c, s):
  //   phi += c k0 + s k1,  L10 += c k2 + s k3,  L11 += (c k4 + s k5, c k6 + s k7)
  // from f = 1 (m = 0) or 2, A = mu_{l+1,m+1}, B = mu_{l+1,m-1} (m >= 1).
  if (t < kNM) {
    int l, m;
    tdecode(t, l, m);
    const double f = m == 0 ? 1.0 : 2.0;
    const double2 u = mus[t];
    double k[8] = {f * u.x, -f * u.y, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (l < kLM) {
      const double2 a = mus[tix(l + 1, m)];
      k[2] = f * a.x;
      k[3] = -f * a.y;
      const double2 A = mus[tix(l + 1, m + 1)];
      const double2 B = m >= 1 ? mus[tix(l + 1, m - 1)] : make_double2(0.0, 0.0);
      const double sb = m == 0 ? 0.0 : (m == 1 ? -1.0 : 1.0);
      k[4] = A.x - B.x;
      k[5] = -A.y + sb * B.y;
      k[6] = A.y + sb * B.y;
      k[7] = A.x + B.x;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) coef[8 * t + j] = k[j];
  }
}

// Lattice field per particle (fmm_sum_lattice_force :766-833):
//   phi = Re L00,  E = -(Re L11, Im L11, Re L10),  L = L2L(O(r - centre), mu)
// with
//   L00 = sum_l O_l0 mu_l0 + 2 Re sum_{m>0} O_lm mu_lm
//   L10 = sum_l O_l0 mu_{l+1,0} + 2 Re sum_{m>0} O_lm mu_{l+1,m}
//   L11 = sum_{m>=0} O_lm mu_{l+1,m+1} - conj(O_l1) mu_{l+1,0}
//         - sum_{m>=2} conj(O_lm mu_{l+1,m-1}),
// all linear in (Re, Im) O_lm, so each entry is eight multiply-adds against
// the coefficients lattice_mu_kernel tabulates; O_lm column by column (order
// m, l = m..kLM by the recurrence).  Then the extrinsic correction
// (:807-829): Redlack-Grindlay adds E += 2 pf D and phi += -2 pf r.D + pf Q
// (pf = 2 pi / 3V, D the box dipole, Q the quadrupole trace, r relative to
// the centre); the fictitious charges of every neighbour box, merged by
// lattice point (plain Coulomb pair, no softening; 1/r by reciprocal square
// root), otherwise, when set.  Coefficients, recurrence constants, box vectors and charges are
// wave-uniform and come through the scalar cache.  Adds sign * (E, phi) to
// the walk's results.
constexpr int kFieldT = 256;

__global__ __launch_bounds__(kFieldT) void lattice_field_kernel(int64_t n, const double* __restrict__ x,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ z,
                                                               LatticeGeom g,
                                                               const double* __restrict__ coef,
                                                               const double* __restrict__ inv,
                                                               const double* __restrict__ fict,
                                                               double sign, double* __restrict__ ex,
                                                               double* __restrict__ ey,
                                                               double* __restrict__ ez,
                                                               double* __restrict__ pot,
                                                               const int64_t* __restrict__ idx) {
#pragma clang fp contract(fast)   // tolerance-checked sums: fused multiply-adds allowed
  const int64_t i = blockIdx.x * (int64_t)kFieldT + threadIdx.x;
  if (i >= n) return;
  const int64_t at = idx ? idx[i] : i;   // position of output i
  const double px = x[at], py = y[at], pz = z[at];
  const double rx = px - g.center[0], ry = py - g.center[1], rz = pz - g.center[2];
  const double r2 = rx * rx + ry * ry + rz * rz;
  double phi = 0.0, l10 = 0.0, l11r = 0.0, l11i = 0.0;
  double wr = 1.0, wi = 0.0;     // O_mm
  for (int m = 0; m <= kLM; ++m) {
    if (m > 0) {
      const double h = inv[kNM + m];          // 1 / (2m)
      const double tr = (wr * rx + wi * ry) * h, ti = (wi * rx - wr * ry) * h;
      wr = tr; wi = ti;
    }
    double cr = wr, ci = wi, p1r = 0.0, p1i = 0.0;
#pragma unroll 4
    for (int l = m; l <= kLM; ++l) {
      if (l > m) {
        const double a = (double)(2 * l - 1) * rz, sc = inv[tix(l, m)];
        const double nr = (a * p1r - r2 * cr) * sc, ni = (a * p1i - r2 * ci) * sc;
        cr = p1r; ci = p1i;           // O_{l-2}
        p1r = nr; p1i = ni;           // O_l
      } else {
        p1r = wr; p1i = wi; cr = 0.0; ci = 0.0;
      }
      const double* k = coef + 8 * tix(l, m);
      phi = fma(p1r, k[0], fma(p1i, k[1], phi));
      l10 = fma(p1r, k[2], fma(p1i, k[3], l10));
      l11r = fma(p1r, k[4], fma(p1i, k[5], l11r));
      l11i = fma(p1r, k[6], fma(p1i, k[7], l11i));
    }
  }
  double e0 = -l11r, e1 = -l11i, e2 = -l10;
  if (g.correction == kExtrinsicRedlack) {
    const double pf = 2.0 * 3.141592653589793 / (3.0 * g.volume);
    const double dx = fict[17], dy = fict[18], dz = fict[19], qt = fict[20];
    e0 += 2.0 * pf * dx;
    e1 += 2.0 * pf * dy;
    e2 += 2.0 * pf * dz;
    phi += -2.0 * pf * (rx * dx + ry * dy + rz * dz) + pf * qt;
  }
  for (int k = 0; k < g.nfpt; ++k) {
    const double* f = g.fpos + 4 * k;
    const double dx = px - f[0], dy = py - f[1], dz = pz - f[2];
    const double rd = rsqrt(dx * dx + dy * dy + dz * dz);
    const double qrd = f[3] * rd;
    const double rd3q = qrd * rd * rd;
    phi += qrd;
    e0 += rd3q * dx; e1 += rd3q * dy; e2 += rd3q * dz;
  }
  ex[i] += sign * e0;
  ey[i] += sign * e1;
  ez[i] += sign * e2;
  pot[i] += sign * phi;
}

// inv[l(l+1)/2 + m] = 1 / ((l-m)(l+m)) for l > m; inv[kNM + m] = 1 / (2m).
__global__ void lattice_inv_kernel(double* __restrict__ inv) {
  for (int t = threadIdx.x; t < kNM + kLM + 1; t += blockDim.x) {
    if (t < kNM) {
      int l, m;
      tdecode(t, l, m);
      inv[t] = l > m ? 1.0 / (double)((l - m) * (l + m)) : 0.0;
    } else {
      const int m = t - kNM;
      inv[t] = m > 0 ? 0.5 / (double)m : 0.0;
    }
  }
}

// constrain_periodic (module_mirror_boxes.f90:174-210): Fortran MODULO,
// a - floor(a / p) p, per periodic axis for a principal-axes lattice; else in
// lattice coordinates (x - origin) Lattice^-1, wrapped to [0, 1) when any
// leaves [0, 1], mapped back, periodic axes replaced.
__global__ void constrain_kernel(int64_t n, double* __restrict__ x, double* __restrict__ y,
                                 double* __restrict__ z, LatticeGeom g) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double p[3] = {x[i], y[i], z[i]};
    if (g.simple) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (!g.per[d]) continue;
        const double w = g.lat[4 * d];
        const double a = p[d] - g.origin[d];
        p[d] = (a - floor(a / w) * w) + g.origin[d];
      }
    } else {
      const double v[3] = {p[0] - g.origin[0], p[1] - g.origin[1], p[2] - g.origin[2]};
      double lc[3];
      bool out = false;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        lc[j] = v[0] * g.latinv[j] + v[1] * g.latinv[3 + j] + v[2] * g.latinv[6 + j];
        out = out || lc[j] > 1.0 || lc[j] < 0.0;
      }
      if (out) {
#pragma unroll
        for (int j = 0; j < 3; ++j) lc[j] = lc[j] - floor(lc[j]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double r = lc[0] * g.lat[d] + lc[1] * g.lat[3 + d] + lc[2] * g.lat[6 + d] + g.origin[d];
          if (g.per[d]) p[d] = r;
        }
      }
    }
    x[i] = p[0]; y[i] = p[1]; z[i] = p[2];
  }
}

}  // namespace

void launch_constrain_periodic(hipStream_t s, int64_t n, double* x, double* y, double* z,
                               const LatticeGeom& g) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  constrain_kernel<<<(unsigned)b, 256, 0, s>>>(n, x, y, z, g);
}

void launch_lattice_coefficients(hipStream_t s, const double* vec, int nb, double scale,
                                 const double* fact, double2* mstar, double2* lstar,
                                 int maxiter, double2* ml) {
  if ((int64_t)(nb - 1) * nb <= kTermCap)   // one mirror layer: sorted sums as the reference
    lattice_terms_kernel<<<kNM + kNT, kTermCap, 0, s>>>(vec, nb, scale, fact, mstar, lstar);
  else
    lattice_terms_dd_kernel<<<kNM + kNT, kTermDD, 0, s>>>(vec, nb, scale, fact, mstar, lstar);
  lattice_iterate_kernel<<<1, 1024, 0, s>>>(mstar, lstar, scale, maxiter, ml);
}

int lattice_moment_blocks(int64_t n) {
  // at least 16 particles per thread, at most 256 blocks per row
  int64_t b = (n + kMomT * 16 - 1) / (kMomT * 16);
  if (b < 1) b = 1;
  if (b > 256) b = 256;
  return (int)b;
}

void launch_lattice_inv(hipStream_t s, double* inv) { lattice_inv_kernel<<<1, 256, 0, s>>>(inv); }

void launch_lattice_sources(hipStream_t s, int64_t n, const double* x, const double* y,
                            const double* z, const double* q, const LatticeGeom& g,
                            const double2* ml, const double* fact, double2* part, double* dpart,
                            double2* om, double* dip, double2* mu, double* fict, double* coef) {
  const int nblk = lattice_moment_blocks(n);
  lattice_moments_kernel<<<dim3(nblk, kLM + 1), kMomT, 0, s>>>(n, x, y, z, q, g, part, dpart);
  lattice_omega_kernel<<<kNM + 4, kWave, 0, s>>>(part, dpart, nblk, om, dip);
  lattice_mu_kernel<<<1, kMuT, 0, s>>>(om, dip, g, ml, fact, mu, fict, coef);
}

void launch_lattice_field(hipStream_t s, int64_t m, const double* x, const double* y, const double* z,
                          const LatticeGeom& g, const double* coef, const double* inv,
                          const double* fict, double sign, double* ex, double* ey, double* ez,
                          double* pot, const int64_t* idx) {
  if (m <= 0) return;
  const unsigned fb = (unsigned)((m + kFieldT - 1) / kFieldT);
  lattice_field_kernel<<<fb, kFieldT, 0, s>>>(m, x, y, z, g, coef, inv, fict, sign, ex, ey, ez,
                                              pot, idx);
}

void launch_lattice_far(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                        const double* q, const LatticeGeom& g, const double2* ml, const double* fact,
                        const double* inv, double sign, double2* part,
                        double* dpart, double2* om, double* dip, double2* mu, double* fict,
                        double* coef, double* ex, double* ey, double* ez, double* pot) {
  launch_lattice_sources(s, n, x, y, z, q, g, ml, fact, part, dpart, om, dip, mu, fict, coef);
  launch_lattice_field(s, n, x, y, z, g, coef, inv, fict, sign, ex, ey, ez, pot);
}

}  // namespace pepc
