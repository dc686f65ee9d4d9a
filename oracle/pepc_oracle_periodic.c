/*
 * pepc_oracle_periodic.c -- CPU restatement of PEPC's periodic boundaries for
 * the Coulomb backend: the mirror boxes of the near field and the lattice
 * (far-field) contribution of the fast-multipole framework.
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so, see pepc_oracle.h).
 *
 * References:
 *   src/treecode/lpepcsrc/module_mirror_boxes.f90   neighbour boxes, lattice
 *   src/interaction_specific/coulomb/module_fmm_framework.f90
 *       lattice coefficients (Challacombe, White and Head-Gordon, J. Chem.
 *       Phys. 107, 10131 (1997)), fictitious-charge extrinsic correction
 *       (Kudin and Scuseria, Chem. Phys. Lett. 283, 61 (1998)), far field.
 * Expansion tables hold l = 0..lmax, m = 0..l at idx(l, m) = l(l+1)/2 + m
 * (tblinv, :863-873); negative orders are (-1)^m conj (tbl, :838-858).
 * Complex values cross the ABI as interleaved (re, im) doubles.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pepc_oracle.h"

typedef double complex cplx;

enum { LM = OR_LMAX_MULT, LT = OR_LMAX_TAYLOR, NM = OR_NMULT, NT = OR_NTAYLOR };

static int tblinv(int l, int m) { return l * (l + 1) / 2 + m; }

/* :1247-1319: the exact integers up to 24!, gamma(n+1) beyond. */
static double factorial(int n)
{
    if (n <= 24) {
        unsigned __int128 f = 1;
        for (int k = 2; k <= n; ++k) f *= (unsigned)k;
        return (double)f;
    }
    return tgamma((double)n + 1.0);
}

/* Associated Legendre function without the Condon-Shortley phase,
 * :1196-1245 (the recurrence carries (-1)^m, undone at the end). */
static double legendre_p(int l, int m, double x)
{
    double pmm = 1.0, pll = 0.0, v;
    if (m > 0) {
        double somx2 = sqrt((1.0 - x) * (1.0 + x)), fact = 1.0;
        for (int i = 1; i <= m; ++i) { pmm = -pmm * fact * somx2; fact += 2.0; }
    }
    if (l == m) {
        v = pmm;
    } else {
        double pmmp1 = x * (2.0 * m + 1.0) * pmm;
        if (l == m + 1) {
            v = pmmp1;
        } else {
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

static double ptilda(int l, int m, double x)      /* :659-672 */
{
    if (m >= 0) return legendre_p(l, m, x) / factorial(l + m);
    return ((m & 1) ? -1.0 : 1.0) * ptilda(l, -m, x);
}

static double p2tilda(int l, int m, double x)     /* :677-690 */
{
    if (m >= 0) return legendre_p(l, m, x) * factorial(l - m);
    return ((m & 1) ? -1.0 : 1.0) * p2tilda(l, -m, x);
}

static void to_spherical(const double c[3], double s[3])   /* :1083-1103 */
{
    s[0] = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    s[1] = s[0] == 0.0 ? 1.0 : c[2] / s[0];
    s[2] = (c[0] == 0.0 && c[1] == 0.0) ? 0.0 : atan2(c[1], c[0]);
}

static cplx cexpi(double a) { return cos(a) + I * sin(a); }

static cplx o_multipole(int l, int m, const double s[3])  /* :695-715 */
{
    if (l < 0 || m < -l || m > l) return 0.0;
    double rl = (s[0] > 0.0 || l > 0) ? pow(s[0], (double)l) : 1.0;
    return rl * ptilda(l, m, s[1]) * cexpi(-(double)m * s[2]);
}

static cplx m_taylor(int l, int m, const double s[3])     /* :720-734 */
{
    if (l < 0 || m < -l || m > l) return 0.0;
    return 1.0 / pow(s[0], (double)(l + 1)) * p2tilda(l, m, s[1]) * cexpi((double)m * s[2]);
}

static cplx tbl(const cplx *a, int l, int m, int lmax)   /* :838-858 */
{
    if (l > lmax || abs(m) > l) return 0.0;
    cplx v = a[tblinv(l, abs(m))];
    if (m < 0) v = ((m & 1) ? -1.0 : 1.0) * conj(v);
    return v;
}

/* M2L (:912-938): out(l,m) = sum_j (-1)^j sum_k Mb(j+l, k+m) Oa(j, k). */
static void m2l(const cplx *mb, const cplx *oa, cplx *out)
{
    for (int l = 0; l <= LT; ++l)
        for (int m = 0; m <= l; ++m) {
            cplx t = 0.0;
            for (int j = 0; j <= LM; ++j) {          /* Oa(j > LM) = 0 */
                double sj = (j & 1) ? -1.0 : 1.0;
                for (int k = -j; k <= j; ++k)
                    t += sj * tbl(mb, j + l, k + m, LT) * tbl(oa, j, k, LM);
            }
            out[tblinv(l, m)] = t;
        }
}

/* L2L (:943-977) up to order maxl. */
static void l2l(const cplx *ob, const cplx *mr, int maxl, cplx *out)
{
    for (int l = 0; l <= maxl; ++l)
        for (int m = 0; m <= l; ++m) {
            cplx t = 0.0;
            for (int j = l; j <= LM; ++j)
                for (int k = -j; k <= j; ++k)
                    t += tbl(ob, j - l, k - m, LM) * tbl(mr, j, k, LT);
            out[tblinv(l, m)] = t;
        }
}

/* module_mirror_boxes.f90:104-151: boxes of the periodic directions within
 * `layers` shells, a box followed by its central-symmetric partner, the
 * central box last.  Returns the count; boxes (3 per box) may be NULL. */
int or_neighbour_boxes(const int32_t per[3], int32_t layers, int32_t *boxes)
{
    int sw[3];
    for (int d = 0; d < 3; ++d) sw[d] = per[d] ? layers : 0;
    int idx = 0;
#define PUT(a, b, c)                                                                   \
    do {                                                                               \
        if (boxes) { boxes[3 * idx] = (a); boxes[3 * idx + 1] = (b); boxes[3 * idx + 2] = (c); } \
        idx++;                                                                         \
    } while (0)
    for (int i = -sw[0]; i <= -1; ++i)
        for (int j = -sw[1]; j <= sw[1]; ++j)
            for (int k = -sw[2]; k <= sw[2]; ++k) { PUT(i, j, k); PUT(-i, -j, -k); }
    for (int j = -sw[1]; j <= -1; ++j)
        for (int k = -sw[2]; k <= sw[2]; ++k) { PUT(0, j, k); PUT(0, -j, -k); }
    for (int k = -sw[2]; k <= -1; ++k) { PUT(0, 0, k); PUT(0, 0, -k); }
    PUT(0, 0, 0);
#undef PUT
    return idx;
}

/* lattice_vect, module_mirror_boxes.f90:73-79 (lat = t_lattice_1..3, row-wise). */
void or_lattice_vect(const double lat[9], const int32_t ijk[3], double v[3])
{
    for (int d = 0; d < 3; ++d)
        v[d] = (double)ijk[0] * lat[d] + (double)ijk[1] * lat[3 + d] + (double)ijk[2] * lat[6 + d];
}

static int cmp_abs(const void *a, const void *b)
{
    double x = fabs(*(const double *)a), y = fabs(*(const double *)b);
    return (x > y) - (x < y);
}

/* Sum of terms from the smallest magnitude up, real and imaginary parts
 * separately (MstarFunc / LstarFunc, :1005-1077). */
static cplx sorted_sum(double *re, double *im, int n)
{
    qsort(re, (size_t)n, sizeof(double), cmp_abs);
    qsort(im, (size_t)n, sizeof(double), cmp_abs);
    double rp = 0.0, ip = 0.0;
    for (int i = 0; i < n; ++i) { rp += re[i]; ip += im[i]; }
    return rp + I * ip;
}

/* calc_lattice_coefficients (:154-213).  ml: 2*OR_NTAYLOR doubles. */
int or_lattice_coefficients(const double lat[9], const int32_t per[3], int32_t layers,
                            int32_t maxiter, double *ml_out)
{
    int nb = or_neighbour_boxes(per, layers, NULL);
    int32_t *boxes = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)nb);
    double *vec = (double *)malloc(sizeof(double) * 3 * (size_t)nb);
    size_t nterm = (size_t)nb * (size_t)(nb - 1) + (size_t)nb;
    double *re = (double *)malloc(sizeof(double) * nterm);
    double *im = (double *)malloc(sizeof(double) * nterm);
    cplx *mstar = (cplx *)calloc(NM, sizeof(cplx));
    cplx *lstar = (cplx *)calloc(NT, sizeof(cplx));
    cplx *ml = (cplx *)calloc(NT, sizeof(cplx));
    cplx *ul = (cplx *)calloc(NT, sizeof(cplx));
    if (!boxes || !vec || !re || !im || !mstar || !lstar || !ml || !ul) return -1;
    or_neighbour_boxes(per, layers, boxes);
    for (int i = 0; i < nb; ++i) or_lattice_vect(lat, boxes + 3 * i, vec + 3 * i);
    const double scale = (double)(2 * layers + 1);

    for (int l = 0; l <= LM; ++l)                       /* MstarFunc */
        for (int m = 0; m <= l; ++m) {
            for (int i = 0; i < nb; ++i) {
                double c[3] = {-vec[3 * i], -vec[3 * i + 1], -vec[3 * i + 2]}, s[3];
                to_spherical(c, s);
                cplx t = o_multipole(l, m, s);
                re[i] = creal(t); im[i] = cimag(t);
            }
            mstar[tblinv(l, m)] = sorted_sum(re, im, nb);
        }
    for (int l = 0; l <= LT; ++l)                       /* LstarFunc */
        for (int m = 0; m <= l; ++m) {
            int k = 0;
            for (int i = 0; i < nb - 1; ++i)
                for (int ii = 0; ii < nb; ++ii) {
                    double c[3], s[3];
                    for (int d = 0; d < 3; ++d) c[d] = scale * vec[3 * i + d] + vec[3 * ii + d];
                    to_spherical(c, s);
                    cplx t = m_taylor(l, m, s);
                    re[k] = creal(t); im[k] = cimag(t); k++;
                }
            lstar[tblinv(l, m)] = sorted_sum(re, im, k);
        }
    memcpy(ml, lstar, sizeof(cplx) * NT);
    for (int it = 0; it < maxiter; ++it) {               /* ML = M2L(UL(ML), M*) + L* */
        for (int l = 0; l <= LT; ++l)                    /* UL, :982-1000 */
            for (int m = 0; m <= l; ++m)
                ul[tblinv(l, m)] = ml[tblinv(l, m)] / pow(scale, (double)(l + 1));
        m2l(ul, mstar, ml);
        for (int i = 0; i < NT; ++i) ml[i] += lstar[i];
    }
    memcpy(ml_out, ml, sizeof(cplx) * NT);
    free(boxes); free(vec); free(re); free(im); free(mstar); free(lstar); free(ml); free(ul);
    return 0;
}

static void lattice_center(const double lat[9], const double origin[3], double c[3])
{
    for (int d = 0; d < 3; ++d)                     /* module_mirror_boxes.f90:106 */
        c[d] = 0.5 * (lat[d] + lat[3 + d] + lat[6 + d]) + origin[d];
}

/* fmm_framework_timestep (:132-149): box dipole and quadrupole trace
 * (calc_box_dipole :605-631), omega_tilde of the particles (and, with the
 * fictitious-charge correction, of the fictitious charges) about the box
 * centre with the monopole removed (calc_omega_tilde :514-600), mu_cent =
 * M2L(ML, omega) (calc_mu_cent :636-654).  correction: OR_EXTRINSIC_NONE,
 * _REDLACK or _FICTCHARGE (module_fmm_framework.f90:42-46).  Outputs: omega
 * (2*OR_NMULT), mu (2*OR_NTAYLOR), fict (4 x {x, y, z, q}; none unless
 * FICTCHARGE), nfict, dipole[3], quad_trace[1]. */
void or_lattice_timestep(const double lat[9], const double origin[3], const int32_t per[3],
                         const double *ml_in, int64_t n, const double *x, const double *y,
                         const double *z, const double *q, int32_t correction, double *omega_out,
                         double *mu_out, double *fict, int32_t *nfict, double dipole[3],
                         double *quad_trace)
{
    double c[3];
    lattice_center(lat, origin, c);
    double dp[3] = {0.0, 0.0, 0.0}, qt = 0.0;
    for (int64_t p = 0; p < n; ++p) {
        double r[3] = {x[p] - c[0], y[p] - c[1], z[p] - c[2]};
        for (int d = 0; d < 3; ++d) dp[d] += q[p] * r[d];
        qt += q[p] * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    }
    cplx *om = (cplx *)calloc(NM, sizeof(cplx));
    for (int64_t p = 0; p < n; ++p) {
        double r[3] = {x[p] - c[0], y[p] - c[1], z[p] - c[2]}, s[3];
        to_spherical(r, s);
        for (int l = 0; l <= LM; ++l)
            for (int m = 0; m <= l; ++m) om[tblinv(l, m)] += q[p] * o_multipole(l, m, s);
    }
    int nf = 0;
    if (correction == OR_EXTRINSIC_FICTCHARGE) {
        double fq_sum = 0.0;
        for (int p = 0; p < 3; ++p) {
            if (!per[p]) continue;
            const double *t = lat + 3 * p;
            for (int d = 0; d < 3; ++d) fict[4 * nf + d] = origin[d] + t[d];
            fict[4 * nf + 3] = -dp[p] / sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            fq_sum += fict[4 * nf + 3];
            nf++;
        }
        for (int d = 0; d < 3; ++d) fict[4 * nf + d] = origin[d];
        fict[4 * nf + 3] = -fq_sum;
        nf++;
    }
    for (int f = 0; f < nf; ++f) {
        double r[3] = {fict[4 * f] - c[0], fict[4 * f + 1] - c[1], fict[4 * f + 2] - c[2]}, s[3];
        to_spherical(r, s);
        for (int l = 0; l <= LM; ++l)
            for (int m = 0; m <= l; ++m) om[tblinv(l, m)] += fict[4 * f + 3] * o_multipole(l, m, s);
    }
    if (cabs(om[0]) > 0.0) om[0] = 0.0;
    cplx *mu = (cplx *)calloc(NT, sizeof(cplx));
    m2l((const cplx *)ml_in, om, mu);
    memcpy(omega_out, om, sizeof(cplx) * NM);
    memcpy(mu_out, mu, sizeof(cplx) * NT);
    for (int d = 0; d < 3; ++d) dipole[d] = dp[d];
    *quad_trace = qt;
    *nfict = nf;
    free(om); free(mu);
}

/* fmm_sum_lattice_force (:766-833) per position: mu shifted to the particle
 * (L2L to order 1), E = -grad, then the extrinsic correction: REDLACK adds
 * E += 2 pf D, phi += -2 pf (r - c).D + pf Q with pf = 2 pi / (3 V), D the
 * box dipole and Q the quadrupole trace (:810-812); FICTCHARGE adds the
 * fictitious charges of every neighbour box (plain Coulomb pair, no
 * softening, :813-823); NONE adds nothing.  V = |(t1 x t2).t3|
 * (module_mirror_boxes.f90:153). */
void or_lattice_field(const double lat[9], const double origin[3], const int32_t per[3],
                      int32_t layers, const double *mu_in, int32_t correction,
                      const double *fict, int32_t nfict, const double dipole[3], double quad_trace,
                      int64_t n, const double *x, const double *y, const double *z,
                      double *ex, double *ey, double *ez, double *pot)
{
    double c[3];
    lattice_center(lat, origin, c);
    const double vol = fabs((lat[1] * lat[5] - lat[2] * lat[4]) * lat[6] +
                            (lat[2] * lat[3] - lat[0] * lat[5]) * lat[7] +
                            (lat[0] * lat[4] - lat[1] * lat[3]) * lat[8]);
    const double pf = 2.0 * acos(-1.0) / (3.0 * vol);
    if (correction != OR_EXTRINSIC_FICTCHARGE) nfict = 0;
    int nb = or_neighbour_boxes(per, layers, NULL);
    int32_t *boxes = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)nb);
    or_neighbour_boxes(per, layers, boxes);
    const cplx *mu = (const cplx *)mu_in;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < n; ++p) {
        cplx o[NM], ms[3];
        double r[3] = {x[p] - c[0], y[p] - c[1], z[p] - c[2]}, s[3];
        to_spherical(r, s);
        for (int l = 0; l <= LM; ++l)
            for (int m = 0; m <= l; ++m) o[tblinv(l, m)] = o_multipole(l, m, s);
        l2l(o, mu, 1, ms);
        double e[3] = {-creal(ms[2]), -cimag(ms[2]), -creal(ms[1])};
        double phi = creal(ms[0]);
        if (correction == OR_EXTRINSIC_REDLACK) {
            for (int d = 0; d < 3; ++d) e[d] += 2.0 * pf * dipole[d];
            phi += -2.0 * pf * (r[0] * dipole[0] + r[1] * dipole[1] + r[2] * dipole[2]) +
                   pf * quad_trace;
        }
        for (int f = 0; f < nfict; ++f)
            for (int b = 0; b < nb; ++b) {
                double v[3], d[3];
                or_lattice_vect(lat, boxes + 3 * b, v);
                d[0] = x[p] - v[0] - fict[4 * f];
                d[1] = y[p] - v[1] - fict[4 * f + 1];
                d[2] = z[p] - v[2] - fict[4 * f + 2];
                double exyz[3], ph;
                or_p2p(&fict[4 * f + 3], d, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], exyz, &ph);
                e[0] += exyz[0]; e[1] += exyz[1]; e[2] += exyz[2];
                phi += ph;
            }
        ex[p] = e[0]; ey[p] = e[1]; ez[p] = e[2]; pot[p] = phi;
    }
    free(boxes);
}
