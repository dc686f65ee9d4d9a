/*
 * pepc_oracle.c -- CPU restatement of PEPC's Coulomb Barnes-Hut force solve.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the CPU
 * baseline of bench.py.  Nothing under pepc_amd/ links or loads this file.
 * Compiled with -O3 -fopenmp -ffp-contract=off (see oracle/Makefile).
 *
 * The functions follow the reference's algorithm step by step (cited per
 * function); the code is this project's own.  Where the reference's operation
 * order fixes floating-point bits (moment shifts, MAC, kernels) the C
 * expressions below are written in the same left-to-right order as the
 * Fortran source.  The per-function parity pins are listed in pepc_oracle.h.
 */
#include "pepc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* box: module_box.f90:48-93                                                 */
/* ------------------------------------------------------------------------- */
int or_box_create(int64_t n, const double *x, const double *y, const double *z,
                  int force_cubic, or_box *b)
{
    if (n < 1) return -1;
    double mn[3] = {x[0], y[0], z[0]}, mx[3] = {x[0], y[0], z[0]};
    const double *c[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        for (int64_t i = 1; i < n; ++i) {
            double v = c[d][i];
            if (v < mn[d]) mn[d] = v;
            if (v > mx[d]) mx[d] = v;
        }
    }
    if (force_cubic) { /* module_box.f90:71-74 */
        double lo = fmin(fmin(mn[0], mn[1]), mn[2]);
        double hi = fmax(fmax(mx[0], mx[1]), mx[2]);
        for (int d = 0; d < 3; ++d) { mn[d] = lo; mx[d] = hi; }
    }
    for (int d = 0; d < 3; ++d) {
        if (!isfinite(mn[d]) || !isfinite(mx[d])) return -2;
        /* module_box.f90:81-84: margin of boxsize/10000 on both sides */
        double size0 = mx[d] - mn[d];
        b->boxmin[d] = mn[d] - size0 / 10000.0;
        b->boxmax[d] = mx[d] + size0 / 10000.0;
        b->boxsize[d] = b->boxmax[d] - b->boxmin[d];
        if (!(b->boxsize[d] > 0.0)) return -3;
    }
    return 0;
}

/* module_tree.f90:196-201: boxlength2(0) = maxval(boxsize)**2, /4 per level */
void or_boxlength2(const or_box *b, double b2[OR_MAXLEVEL + 1])
{
    double m = b->boxsize[0];
    if (b->boxsize[1] > m) m = b->boxsize[1];
    if (b->boxsize[2] > m) m = b->boxsize[2];
    b2[0] = m * m;
    for (int l = 1; l <= OR_MAXLEVEL; ++l) b2[l] = b2[l - 1] / 4.0;
}

/* ------------------------------------------------------------------------- */
/* Hilbert keys: module_spacefilling.f90:224-257 (compute_particle_keys),     */
/* :208-220 (coord_to_intcoord), :380-440 (intcoord_to_key_hilbert3D).       */
/* Kamata et al. 1999 tables, as the reference states them at :389-390.       */
/* ------------------------------------------------------------------------- */
static const int CI3[8] = {0, 1, 3, 2, 7, 6, 4, 5};   /* inverse Hilbert cell */
static const int C3[8] = {0, 1, 3, 2, 6, 7, 5, 4};    /* Hilbert cell (:534) */
static const int G3X[8] = {5, 6, 0, 5, 5, 0, 6, 5};   /* gene: exchange */
static const int G3R[8] = {0, 0, 0, 5, 0, 0, 6, 5};   /* gene: reverse  */

int or_level_from_key(uint64_t key)
{
    /* module_spacefilling.f90:47-57: (bit_size - leadz - 1)/idim */
    if (key == 0) return -1;
    return (63 - __builtin_clzll(key)) / 3;
}

uint64_t or_intcoord_to_key_hilbert3d(int64_t ix, int64_t iy, int64_t iz)
{
    int64_t it[3] = {ix, iy, iz};
    uint64_t key = 1;                 /* placeholder bit */
    for (int i = OR_MAXLEVEL - 1; i >= 0; --i) {
        uint64_t cval = 0;
        for (int j = 2; j >= 0; --j) { /* z, y, x -> MSB..LSB */
            cval = (cval << 1) | (uint64_t)((it[j] >> i) & 1);
            it[j] &= ~((int64_t)1 << i);
        }
        int horder = CI3[cval];
        int exch = G3X[horder];
        int rev = G3R[horder];
        key = (key << 3) | (uint64_t)horder;
        if (exch == 5) { int64_t t = it[0]; it[0] = it[2]; it[2] = t; }
        else if (exch == 6) { int64_t t = it[1]; it[1] = it[2]; it[2] = t; }
        for (int j = 0; j < 3; ++j)
            if ((rev >> j) & 1) it[j] = (~it[j]) & (((int64_t)1 << i) - 1);
    }
    return key;
}

void or_compute_keys(int64_t n, const double *x, const double *y, const double *z,
                     const or_box *b, uint64_t *keys)
{
    double s[3];
    for (int d = 0; d < 3; ++d) s[d] = b->boxsize[d] / 1048576.0; /* 2**maxlevel */
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        int64_t ix = (int64_t)((x[j] - b->boxmin[0]) / s[0]);
        int64_t iy = (int64_t)((y[j] - b->boxmin[1]) / s[1]);
        int64_t iz = (int64_t)((z[j] - b->boxmin[2]) / s[2]);
        keys[j] = or_intcoord_to_key_hilbert3d(ix, iy, iz);
    }
}

/* key_to_intcoord_hilbert, module_spacefilling.f90:520-565 */
void or_key_to_intcoord_hilbert(uint64_t key, int64_t ic[3])
{
    int lev = or_level_from_key(key);
    ic[0] = ic[1] = ic[2] = 0;
    for (int i = 0; i < lev; ++i) {
        int horder = (int)((key >> (3 * i)) & 7);
        int cval = C3[horder];
        int rev = G3R[horder];
        int exch = G3X[horder];
        if (i > 0) {
            for (int j = 0; j < 3; ++j)
                if ((rev >> j) & 1) ic[j] = (~ic[j]) & (((int64_t)1 << i) - 1);
            for (int j = 0; j < 2; ++j) {
                if (!((exch >> j) & 1)) continue;
                for (int k = j + 1; k < 3; ++k)
                    if ((exch >> k) & 1) { int64_t t = ic[j]; ic[j] = ic[k]; ic[k] = t; }
            }
        }
        for (int j = 0; j < 3; ++j) ic[j] |= (int64_t)((cval >> j) & 1) << i;
    }
    for (int j = 0; j < 3; ++j) ic[j] <<= (OR_MAXLEVEL - lev);
}

/* key_to_coord, module_spacefilling.f90:445-471: cell centre */
void or_key_to_coord(const or_box *b, uint64_t key, double xyz[3])
{
    int64_t ic[3];
    or_key_to_intcoord_hilbert(key, ic);
    for (int d = 0; d < 3; ++d) {
        double s = b->boxsize[d] / 1048576.0;
        xyz[d] = ((double)ic[d] + 0.5) * s + b->boxmin[d];
    }
}

/* ------------------------------------------------------------------------- */
/* Local presort: sl_pepc sort_radix.c:33-151 (rs_rec_ip: in-place MSD radix, */
/* 10-bit digits from bit 63 down, cycle-leader class permutation, recursion */
/* above 256 elements) and sort_insert.c:14-37 (masked insertion sort).       */
/* Parameters: sl_tune_intern.h:22-30, mpi_pepckeys.c:29-31, sl_key.h:116.    */
/* The order of equal keys produced here is the reference's order.           */
/* ------------------------------------------------------------------------- */
#define SR_WIDTH 10
#define SR_THRESHOLD 256

typedef struct { uint64_t *k; int64_t *ix; double *w; } sr_elems;

static inline void sr_swap(sr_elems *e, int64_t a, int64_t b)
{
    uint64_t tk = e->k[a]; e->k[a] = e->k[b]; e->k[b] = tk;
    int64_t ti = e->ix[a]; e->ix[a] = e->ix[b]; e->ix[b] = ti;
    if (e->w) { double tw = e->w[a]; e->w[a] = e->w[b]; e->w[b] = tw; }
}

static void sr_insertion(sr_elems *e, int64_t off, int64_t n, uint64_t bmask)
{
    uint64_t *k = e->k + off; int64_t *ix = e->ix + off; double *w = e->w ? e->w + off : NULL;
    for (int64_t i = 1; i < n; ++i) {
        if ((k[i] & bmask) < (k[i - 1] & bmask)) {
            uint64_t tk = k[i]; int64_t ti = ix[i]; double tw = w ? w[i] : 0.0;
            int64_t j = i - 1;
            for (;;) {
                k[j + 1] = k[j]; ix[j + 1] = ix[j]; if (w) w[j + 1] = w[j];
                if (--j < 0) break;
                if (!((tk & bmask) < (k[j] & bmask))) break;
            }
            k[j + 1] = tk; ix[j + 1] = ti; if (w) w[j + 1] = tw;
        }
    }
}

static void sr_rec(sr_elems *e, int64_t off, int64_t n, int rhigh, int rlow, int rwidth)
{
    int cw = rwidth < (rhigh - rlow + 1) ? rwidth : (rhigh - rlow + 1);
    rhigh -= cw - 1;
    int64_t ncl = (int64_t)1 << cw;
    uint64_t mask = (uint64_t)ncl - 1;
    int64_t c[1 << SR_WIDTH], parts[1 << SR_WIDTH];
    for (int64_t i = 0; i < ncl; ++i) c[i] = 0;
    for (int64_t i = 0; i < n; ++i) ++c[(e->k[off + i] >> rhigh) & mask];
    parts[0] = off;
    for (int64_t i = 1; i < ncl; ++i) parts[i] = parts[i - 1] + c[i - 1];
    int64_t end = off;
    for (int64_t i = 0; i < ncl; ++i) {
        end += c[i];
        int64_t xi = parts[i];
        while (xi < end) {
            int64_t j = (int64_t)((e->k[xi] >> rhigh) & mask);
            while (j != i) {
                int64_t kk = (int64_t)((e->k[parts[j]] >> rhigh) & mask);
                if (kk != j) sr_swap(e, xi, parts[j]);
                ++parts[j];
                j = kk;
            }
            ++xi;
        }
    }
    --rhigh;
    if (rhigh >= rlow) {
        uint64_t bmask = 0;
        if (rhigh - rlow + 1 <= 63) bmask = (uint64_t)1 << (rhigh - rlow + 1);
        bmask = (bmask - 1) << rlow;
        int64_t o = off;
        for (int64_t i = 0; i < ncl; ++i) {
            if (c[i] > SR_THRESHOLD) sr_rec(e, o, c[i], rhigh, rlow, rwidth);
            else if (c[i] > 1) sr_insertion(e, o, c[i], bmask);
            o += c[i];
        }
    }
}

void or_sort_radix_ip(int64_t n, uint64_t *keys, int64_t *idx, double *work)
{
    if (n < 2) return;
    sr_elems e = {keys, idx, work};
    sr_rec(&e, 0, n, 63, 0, SR_WIDTH);
}

/* ------------------------------------------------------------------------- */
/* Duplicate-key fix-up of the single-rank decomposition:                    */
/* module_domains.f90:248-307 (key_diffs, nearest-gap search, key_to_coord).  */
/* 0-based restatement of the 1-based loop; returns #particles moved, or <0   */
/* if the key range cannot hold n distinct keys (:254-262).                  */
/* ------------------------------------------------------------------------- */
int64_t or_dupfix(int64_t n, uint64_t *keys, double *x, double *y, double *z,
                  const or_box *b)
{
    if (n < 2) return 0;
    if ((int64_t)(keys[n - 1] - keys[0]) + 1 < n) return -1;
    int64_t *kd = (int64_t *)malloc(sizeof(int64_t) * n);
    uint64_t *lk = (uint64_t *)malloc(sizeof(uint64_t) * n);
    memcpy(lk, keys, sizeof(uint64_t) * n);
    for (int64_t i = 0; i + 1 < n; ++i) kd[i] = (int64_t)(lk[i + 1] - lk[i]);
    kd[n - 1] = 1; /* last rank: no right neighbour */
    for (int64_t i1 = 1; i1 <= n; ++i1) {            /* 1-based as in the source */
        if (kd[i1 - 1] < 1) {
            int64_t jmax = (n - i1) > i1 ? (n - i1) : i1;
            for (int64_t j = 1; j <= jmax; ++j) {
                int64_t dn = i1 - j < 1 ? 1 : i1 - j;
                int64_t up = i1 + j < n - 1 ? i1 + j : n - 1;
                /* The clamped indices (dn, up) only ever re-test an entry that
                 * already failed for a smaller j, so the shifted ranges below
                 * stay inside 1..n; the bounds are asserted for memory safety. */
                if (kd[dn - 1] > 1 && i1 - j >= 1) {
                    for (int64_t t = i1 - j + 1; t <= i1; ++t) lk[t - 1] -= 1;
                    kd[i1 - 1] += 1;
                    kd[i1 - j - 1] -= 1;
                    break;
                } else if (kd[up - 1] > 1 && i1 + j <= n - 1) {
                    for (int64_t t = i1 + 1; t <= i1 + j; ++t) lk[t - 1] += 1;
                    kd[i1 - 1] += 1;
                    kd[i1 + j - 1] -= 1;
                    break;
                }
            }
        }
    }
    int64_t moved = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (lk[i] != keys[i]) {
            double c[3];
            keys[i] = lk[i];
            or_key_to_coord(b, keys[i], c);
            x[i] = c[0]; y[i] = c[1]; z[i] = c[2];
            ++moved;
        }
    }
    free(kd); free(lk);
    return moved;
}

/* domain_decompose, single rank: module_domains.f90:111-367.  Sorting with the
 * presort (irnkl = identity on one rank, mpi_pepckeys.c:352-359), permute,
 * duplicate fix-up.  perm[i] = original 0-based index of sorted particle i. */
int or_domain_decompose(int64_t n, const double *x, const double *y, const double *z,
                        const double *q, int force_cubic,
                        or_box *b, uint64_t *keys_sorted, int64_t *perm,
                        double *xs, double *ys, double *zs, double *qs,
                        int64_t *nfixed)
{
    int rc = or_box_create(n, x, y, z, force_cubic, b);
    if (rc) return rc;
    or_compute_keys(n, x, y, z, b, keys_sorted);
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    or_sort_radix_ip(n, keys_sorted, perm, NULL);
    for (int64_t i = 0; i < n; ++i) {
        int64_t p = perm[i];
        xs[i] = x[p]; ys[i] = y[p]; zs[i] = z[p]; qs[i] = q[p];
    }
    int64_t moved = or_dupfix(n, keys_sorted, xs, ys, zs, b);
    if (moved < 0) return -4;
    if (nfixed) *nfixed = moved;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Tree build: module_tree_grow.f90:552-697 (insert_helper), :702-734 (leaf),  */
/* :762-821 (update from children); module_tree.f90:409-433 (child wiring);   */
/* coulomb/module_interaction_specific.f90:67-79 (multipole_from_particle),    */
/* :84-140 (shift_multipoles_up).  Nodes are numbered in the order the        */
/* reference provisions them (pre-order).                                    */
/* ------------------------------------------------------------------------- */
typedef struct {
    int64_t n, cap, nn, nleaf, ntwig;
    const uint64_t *keys;
    const double *x, *y, *z, *q;
    uint64_t *nkey; int32_t *nlevel; int64_t *nleaves, *nfirst, *nfirst_child, *nnext;
    int32_t *nnchild; double *mom; int64_t *node_leaf;
    int err;
} tb_ctx;

static void shift_multipoles_up(double *par, const double *const *ch, int nch)
{
    /* coulomb/module_interaction_specific.f90:84-140, same operation order */
    double charge = 0.0, absq = 0.0;
    for (int j = 0; j < nch; ++j) charge = charge + ch[j][OR_CHARGE];
    for (int j = 0; j < nch; ++j) absq = absq + ch[j][OR_ABSCHARGE];
    double coc[3] = {0.0, 0.0, 0.0};
    if (absq != 0.0) {
        for (int j = 0; j < nch; ++j)
            for (int d = 0; d < 3; ++d) coc[d] = coc[d] + (ch[j][OR_COCX + d] * ch[j][OR_ABSCHARGE]);
        for (int d = 0; d < 3; ++d) coc[d] = coc[d] / absq;
    } else {
        for (int j = 0; j < nch; ++j)
            for (int d = 0; d < 3; ++d) coc[d] = coc[d] + ch[j][OR_COCX + d];
        for (int d = 0; d < 3; ++d) coc[d] = coc[d] / (double)nch;
    }
    double dip[3] = {0, 0, 0}, quad[3] = {0, 0, 0}, xy = 0, yz = 0, zx = 0;
    for (int j = 0; j < nch; ++j) {
        const double *c = ch[j];
        double s[3];
        for (int d = 0; d < 3; ++d) s[d] = coc[d] - c[OR_COCX + d];
        for (int d = 0; d < 3; ++d) dip[d] = dip[d] + c[OR_DIPX + d] + c[OR_CHARGE] * s[d];
        for (int d = 0; d < 3; ++d)
            quad[d] = quad[d] + c[OR_QUADX + d] + c[OR_DIPX + d] * s[d] + 0.5 * c[OR_CHARGE] * (s[d] * s[d]);
        xy = xy + c[OR_XYQUAD] + c[OR_DIPX] * s[1] + c[OR_DIPY] * s[0] + c[OR_CHARGE] * s[0] * s[1];
        yz = yz + c[OR_YZQUAD] + c[OR_DIPY] * s[2] + c[OR_DIPZ] * s[1] + c[OR_CHARGE] * s[1] * s[2];
        zx = zx + c[OR_ZXQUAD] + c[OR_DIPZ] * s[0] + c[OR_DIPX] * s[2] + c[OR_CHARGE] * s[2] * s[0];
    }
    double bmax = 0.0;
    for (int j = 0; j < nch; ++j) {
        const double *c = ch[j];
        double a = coc[0] - c[OR_COCX], bb = coc[1] - c[OR_COCY], cc = coc[2] - c[OR_COCZ];
        double v = sqrt(a * a + bb * bb + cc * cc) + c[OR_BMAX];
        if (j == 0 || v > bmax) bmax = v;
    }
    memset(par, 0, sizeof(double) * OR_MOM);
    par[OR_COCX] = coc[0]; par[OR_COCY] = coc[1]; par[OR_COCZ] = coc[2];
    par[OR_CHARGE] = charge; par[OR_ABSCHARGE] = absq;
    for (int d = 0; d < 3; ++d) { par[OR_DIPX + d] = dip[d]; par[OR_QUADX + d] = quad[d]; }
    par[OR_XYQUAD] = xy; par[OR_YZQUAD] = yz; par[OR_ZXQUAD] = zx; par[OR_BMAX] = bmax;
}

static int64_t tb_provision(tb_ctx *c)
{
    if (c->nn >= c->cap) { c->err = -2; return -1; }
    return c->nn++;
}

static int64_t insert_helper(tb_ctx *c, uint64_t k, int l, int64_t *cur)
{
    int sh = 3 * (OR_MAXLEVEL - l);
    int first = (*cur < c->n) && ((c->keys[*cur] >> sh) == k);
    int second = (*cur + 1 < c->n) && ((c->keys[*cur + 1] >> sh) == k);
    if (!first && !second) return -1;
    if (first && !second) {           /* leaf: module_tree_grow.f90:654-661 */
        int64_t id = tb_provision(c);
        if (id < 0) return -1;
        int64_t p = *cur;
        c->nkey[id] = k; c->nlevel[id] = l; c->nleaves[id] = 1; c->nfirst[id] = p;
        c->nfirst_child[id] = -1; c->nnext[id] = -1; c->nnchild[id] = 0;
        double *m = c->mom + OR_MOM * id;
        memset(m, 0, sizeof(double) * OR_MOM);
        m[OR_COCX] = c->x[p]; m[OR_COCY] = c->y[p]; m[OR_COCZ] = c->z[p];
        m[OR_CHARGE] = c->q[p]; m[OR_ABSCHARGE] = fabs(c->q[p]);
        c->node_leaf[p] = id;
        c->nleaf++;
        (*cur)++;
        return id;
    }
    if (l >= OR_MAXLEVEL) { c->err = -3; return -1; } /* :663-668 no more levels */
    int64_t id = tb_provision(c);
    if (id < 0) return -1;
    int64_t kids[8]; int nch = 0;
    for (int ic = 0; ic < 8; ++ic) {
        int64_t ch = insert_helper(c, (k << 3) + (uint64_t)ic, l + 1, cur);
        if (c->err) return -1;
        if (ch >= 0) kids[nch++] = ch;
    }
    const double *cm[8];
    int64_t leaves = 0;
    for (int j = 0; j < nch; ++j) { cm[j] = c->mom + OR_MOM * kids[j]; leaves += c->nleaves[kids[j]]; }
    shift_multipoles_up(c->mom + OR_MOM * id, cm, nch);
    c->nkey[id] = k; c->nlevel[id] = l; c->nleaves[id] = leaves; c->nfirst[id] = c->nfirst[kids[0]];
    c->nfirst_child[id] = kids[0]; c->nnext[id] = -1; c->nnchild[id] = nch;
    for (int j = 0; j + 1 < nch; ++j) c->nnext[kids[j]] = kids[j + 1];
    c->nnext[kids[nch - 1]] = -1;
    c->ntwig++;
    return id;
}

int64_t or_tree_build(int64_t n, const uint64_t *keys, const double *x, const double *y,
                      const double *z, const double *q, int64_t cap,
                      uint64_t *nkey, int32_t *nlevel, int64_t *nleaves,
                      int64_t *nfirst, int64_t *nfirst_child, int64_t *nnext_sibling,
                      int32_t *nnchild, double *nmom, int64_t *node_leaf,
                      int64_t *counts_out)
{
    tb_ctx c = {n, cap, 0, 0, 0, keys, x, y, z, q, nkey, nlevel, nleaves, nfirst,
                nfirst_child, nnext_sibling, nnchild, nmom, node_leaf, 0};
    int64_t cur = 0;
    insert_helper(&c, 1, 0, &cur);
    if (c.err) return c.err;
    if (cur != n) return -5;
    if (counts_out) { counts_out[0] = c.nn; counts_out[1] = c.nleaf; counts_out[2] = c.ntwig; }
    return c.nn;
}

/* ------------------------------------------------------------------------- */
/* Interaction kernels: coulomb/module_coulomb_kernels.f90:60-145 (M2P) and   */
/* :242-257 (P2P), written in the Fortran source's evaluation order.         */
/* ------------------------------------------------------------------------- */
void or_p2p(const double *t_charge, const double d[3], double dist2, double exyz[3], double *phi)
{
    double r = sqrt(dist2);
    double rd = 1.0 / r;
    double rd3charge = *t_charge * rd * rd * rd;
    *phi = *t_charge * rd;
    exyz[0] = rd3charge * d[0];
    exyz[1] = rd3charge * d[1];
    exyz[2] = rd3charge * d[2];
}

void or_m2p(const double *t, const double d[3], double dist2, double exyz[3], double *phi)
{
    const double dx = d[0], dy = d[1], dz = d[2];
    const double q1 = t[OR_QUADX], q2 = t[OR_QUADY], q3 = t[OR_QUADZ];
    const double p1 = t[OR_DIPX], p2 = t[OR_DIPY], p3 = t[OR_DIPZ];
    const double xy = t[OR_XYQUAD], yz = t[OR_YZQUAD], zx = t[OR_ZXQUAD], ch = t[OR_CHARGE];
    double r = sqrt(dist2);
    double rd = 1.0 / r;
    double rd2 = rd * rd, rd3 = rd * rd2, rd5 = rd3 * rd2, rd7 = rd5 * rd2;
    double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    double fdx = 3.0 * dx2 * rd5 - rd3;
    double fdy = 3.0 * dy2 * rd5 - rd3;
    double fdz = 3.0 * dz2 * rd5 - rd3;
    double fdxy = 3.0 * dx * dy * rd5;
    double fdyz = 3.0 * dy * dz * rd5;
    double fdxz = 3.0 * dx * dz * rd5;
    double m2rd5 = 2.0 * rd5;
    double m5rd7 = 5.0 * rd7;
    double pre1 = m5rd7 * dx * dy * dz;
    double pre2x = m5rd7 * dx2 - rd5;
    double pre2y = m5rd7 * dy2 - rd5;
    double pre2z = m5rd7 * dz2 - rd5;
    double preQx = pre2x * q1, preQy = pre2y * q2, preQz = pre2z * q3;
    *phi = ch * rd - (dx * p1 + dy * p2 + dz * p3) * rd3
           + fdx * q1 + fdy * q2 + fdz * q3
           + fdxy * xy + fdyz * yz + fdxz * zx;
    exyz[0] = ch * dx * rd3 - (fdx * p1 + fdxy * p2 + fdxz * p3)
              + 3.0 * (dx * ((pre2x - m2rd5) * q1 + preQy + preQz)
                       + dy * pre2x * xy + dz * pre2x * zx + pre1 * yz);
    exyz[1] = ch * dy * rd3 - (fdy * p2 + fdxy * p1 + fdyz * p3)
              + 3.0 * (dy * ((pre2y - m2rd5) * q2 + preQx + preQz)
                       + dx * pre2y * xy + dz * pre2y * yz + pre1 * zx);
    exyz[2] = ch * dz * rd3 - (fdz * p3 + fdyz * p2 + fdxz * p1)
              + 3.0 * (dz * (+(pre2z - m2rd5) * q3 + preQy + preQx)
                       + dx * pre2z * zx + dy * pre2z * yz + pre1 * xy);
}

/* ------------------------------------------------------------------------- */
/* MAC selection, coulomb/module_interaction_specific.f90:256-275:            */
/*   0: Barnes-Hut, accept iff theta2 * dist2 > boxlength2(level);            */
/*   1: Bmax, accept iff theta2 * dist2 > min(bmax**2, 3.0 * boxlength2);     */
/*   any other value: never accept (the reference's N^2 default case).        */
/* ------------------------------------------------------------------------- */
static int g_mac_select = 0;

void or_set_mac_select(int mac_select) { g_mac_select = mac_select; }
int or_get_mac_select(void) { return g_mac_select; }

static inline int or_mac(double theta2, double dist2, double boxlength2, const double *m)
{
    if (g_mac_select == 0) return theta2 * dist2 > boxlength2;
    if (g_mac_select == 1) {
        double bm = m[OR_BMAX] * m[OR_BMAX], c = 3.0 * boxlength2;
        return theta2 * dist2 > (bm < c ? bm : c);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Walk tiles: module_tree_walk_simple.f90:135-212 (maximal key cells holding */
/* at most TILE_SIZE particles; TILE_SIZE=16 at :40 is the reference rule).   */
/* ------------------------------------------------------------------------- */
/* pack > 0 (this project's packed rule, no reference counterpart): runs of
 * consecutive sibling tiles of one parent are merged while the merged count
 * stays <= pack; a sibling that is split further ends the run.  pack = 0 is
 * the reference rule. */
static void tiles_aux(const uint64_t *keys, int64_t lo, int64_t hi, uint64_t k, int l,
                      int64_t G, int64_t pack, int64_t *tf, int64_t *tc, int64_t *nt, int *err)
{
    int64_t np = hi - lo;
    if (np == 0) return;
    if (np <= G) { tf[*nt] = lo; tc[*nt] = np; (*nt)++; return; }
    if (l >= OR_MAXLEVEL) { *err = 1; return; }
    int sh = 3 * (OR_MAXLEVEL - (l + 1));
    int64_t ps = lo, run_first = -1, run_cnt = 0;
    for (int ic = 0; ic < 8; ++ic) {
        uint64_t ck = (k << 3) + (uint64_t)ic;
        int64_t pe = ps;
        while (pe < hi && (keys[pe] >> sh) == ck) ++pe;
        int64_t cc = pe - ps;
        if (cc > 0 && cc <= G) {
            if (pack > 0 && run_first >= 0 && run_cnt + cc <= pack) {
                run_cnt += cc;
            } else {
                if (run_first >= 0) { tf[*nt] = run_first; tc[*nt] = run_cnt; (*nt)++; }
                run_first = ps; run_cnt = cc;
            }
        } else if (cc > G) {
            if (run_first >= 0) { tf[*nt] = run_first; tc[*nt] = run_cnt; (*nt)++; }
            run_first = -1; run_cnt = 0;
            tiles_aux(keys, ps, pe, ck, l + 1, G, pack, tf, tc, nt, err);
        }
        ps = pe;
    }
    if (run_first >= 0) { tf[*nt] = run_first; tc[*nt] = run_cnt; (*nt)++; }
}

int64_t or_walk_tiles(int64_t n, const uint64_t *keys, int64_t tile_size, int64_t pack,
                      int64_t *tile_first, int64_t *tile_count)
{
    int64_t nt = 0; int err = 0;
    tiles_aux(keys, 0, n, 1, 0, tile_size, pack, tile_first, tile_count, &nt, &err);
    return err ? -1 : nt;
}

/* ------------------------------------------------------------------------- */
/* Group walk: module_tree_walk_simple.f90:256-384.  DFS from the root with   */
/* children in key order; leaf -> P2P for every tile particle (self when the  */
/* separation is zero); twig -> MAC per particle in tile order, the first     */
/* failure opens the node for the whole tile; all pass -> M2P for all.        */
/* MAC: coulomb/module_interaction_specific.f90:256-275 (or_mac above).       */
/* ------------------------------------------------------------------------- */
typedef struct {
    const double *x, *y, *z;
    const int32_t *nlevel; const int64_t *nleaves, *nfc, *nns; const double *mom;
    const double *b2; double theta2, eps2, sign;
    double vb[3];             /* lattice vector of the box walked (vbox) */
    int64_t t0, np;
    double e[64][3], pot[64], work[64]; int32_t p2p[64], m2p[64];
    int64_t num_int, num_mac, ni;
} wk_ctx;

static void walk_aux(wk_ctx *w, int64_t node)
{
    const double *m = w->mom + OR_MOM * node;
    double d[64][3], d2[64];
    if (w->nfc[node] < 0) { /* leaf */
        w->ni += 1;
        for (int64_t ip = 0; ip < w->np; ++ip) {
            int64_t p = w->t0 + ip;
            d[ip][0] = (w->x[p] - w->vb[0]) - m[OR_COCX];
            d[ip][1] = (w->y[p] - w->vb[1]) - m[OR_COCY];
            d[ip][2] = (w->z[p] - w->vb[2]) - m[OR_COCZ];
            d2[ip] = d[ip][0] * d[ip][0] + d[ip][1] * d[ip][1] + d[ip][2] * d[ip][2];
            w->num_int++;
            w->work[ip] += 1.0;
            w->p2p[ip]++;
            if (d2[ip] > 0.0) {
                double exyz[3], phi;
                or_p2p(&m[OR_CHARGE], d[ip], d2[ip] + w->eps2, exyz, &phi);
                w->e[ip][0] += w->sign * exyz[0];
                w->e[ip][1] += w->sign * exyz[1];
                w->e[ip][2] += w->sign * exyz[2];
                w->pot[ip] += w->sign * phi;
            }
        }
        return;
    }
    for (int64_t ip = 0; ip < w->np; ++ip) {
        int64_t p = w->t0 + ip;
        w->num_mac++;
        d[ip][0] = (w->x[p] - w->vb[0]) - m[OR_COCX];
        d[ip][1] = (w->y[p] - w->vb[1]) - m[OR_COCY];
        d[ip][2] = (w->z[p] - w->vb[2]) - m[OR_COCZ];
        d2[ip] = d[ip][0] * d[ip][0] + d[ip][1] * d[ip][1] + d[ip][2] * d[ip][2];
        if (!or_mac(w->theta2, d2[ip], w->b2[w->nlevel[node]], m)) {
            for (int64_t c = w->nfc[node]; c >= 0; c = w->nns[c]) walk_aux(w, c);
            return;
        }
    }
    w->ni += w->nleaves[node];
    for (int64_t ip = 0; ip < w->np; ++ip) {
        double exyz[3], phi;
        w->num_int++;
        w->work[ip] += 1.0;
        w->m2p[ip]++;
        or_m2p(m, d[ip], d2[ip] + w->eps2, exyz, &phi);
        w->e[ip][0] += w->sign * exyz[0];
        w->e[ip][1] += w->sign * exyz[1];
        w->e[ip][2] += w->sign * exyz[2];
        w->pot[ip] += w->sign * phi;
    }
}

int or_walk_group(int64_t n, const double *x, const double *y, const double *z,
                  int64_t nnodes, const int32_t *nlevel, const int64_t *nleaves,
                  const int64_t *nfirst_child, const int64_t *nnext_sibling,
                  const double *nmom, const double b2[OR_MAXLEVEL + 1],
                  double theta2, double eps2, double sign,
                  int64_t ntiles, const int64_t *tile_first, const int64_t *tile_count,
                  int64_t tile_stride, int nthreads,
                  double *ex, double *ey, double *ez, double *pot, double *work,
                  int32_t *cnt_p2p, int32_t *cnt_m2p, int64_t *tile_mac,
                  int64_t *totals)
{
    static const double central[3] = {0.0, 0.0, 0.0};
    return or_walk_group_boxes(n, x, y, z, nnodes, nlevel, nleaves, nfirst_child, nnext_sibling,
                               nmom, b2, theta2, eps2, sign, ntiles, tile_first, tile_count,
                               tile_stride, nthreads, 1, central, ex, ey, ez, pot, work, cnt_p2p,
                               cnt_m2p, tile_mac, totals);
}

/* Walks over a list of boxes (module_libpepc_main.f90:184-187: one
 * tree_walk_run per neighbour box, results summed): every tile walks the tree
 * once per lattice vector, with the targets shifted by -vbox. */
int or_walk_group_boxes(int64_t n, const double *x, const double *y, const double *z,
                        int64_t nnodes, const int32_t *nlevel, const int64_t *nleaves,
                        const int64_t *nfirst_child, const int64_t *nnext_sibling,
                        const double *nmom, const double b2[OR_MAXLEVEL + 1],
                        double theta2, double eps2, double sign,
                        int64_t ntiles, const int64_t *tile_first, const int64_t *tile_count,
                        int64_t tile_stride, int nthreads, int64_t nbox, const double *vbox,
                        double *ex, double *ey, double *ez, double *pot, double *work,
                        int32_t *cnt_p2p, int32_t *cnt_m2p, int64_t *tile_mac,
                        int64_t *totals)
{
    (void)nnodes;
    if (tile_stride < 1) tile_stride = 1;
    int64_t tot_int = 0, tot_mac = 0, tot_tiles = 0, bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : tot_int, tot_mac, tot_tiles, bad)
    for (int64_t t = 0; t < ntiles; t += tile_stride) {
        wk_ctx *w = (wk_ctx *)malloc(sizeof(wk_ctx));
        w->x = x; w->y = y; w->z = z; w->nlevel = nlevel; w->nleaves = nleaves;
        w->nfc = nfirst_child; w->nns = nnext_sibling; w->mom = nmom; w->b2 = b2;
        w->theta2 = theta2; w->eps2 = eps2; w->sign = sign;
        w->t0 = tile_first[t]; w->np = tile_count[t];
        w->num_int = 0; w->num_mac = 0; w->ni = 0;
        if (w->np > 64 || w->np < 1) { bad++; free(w); continue; }
        for (int64_t ip = 0; ip < w->np; ++ip) {
            w->e[ip][0] = w->e[ip][1] = w->e[ip][2] = 0.0; w->pot[ip] = 0.0;
            w->work[ip] = 1.0; /* module_domains.f90:318 resets work to 1 */
            w->p2p[ip] = 0; w->m2p[ip] = 0;
        }
        for (int64_t b = 0; b < nbox; ++b) {
            w->vb[0] = vbox[3 * b]; w->vb[1] = vbox[3 * b + 1]; w->vb[2] = vbox[3 * b + 2];
            w->ni = 0;
            walk_aux(w, 0);
            if (w->ni != n) bad++; /* module_tree_walk_simple.f90:287 */
        }
        for (int64_t ip = 0; ip < w->np; ++ip) {
            int64_t p = w->t0 + ip;
            ex[p] = w->e[ip][0]; ey[p] = w->e[ip][1]; ez[p] = w->e[ip][2]; pot[p] = w->pot[ip];
            work[p] = w->work[ip]; cnt_p2p[p] = w->p2p[ip]; cnt_m2p[p] = w->m2p[ip];
        }
        if (tile_mac) tile_mac[t] = w->num_mac;
        tot_int += w->num_int; tot_mac += w->num_mac; tot_tiles++;
        free(w);
    }
    if (totals) { totals[0] = tot_int; totals[1] = tot_mac; totals[2] = tot_tiles; }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* FP64 direct sum: src/utils/module_directsum.f90:34-172 (leaf kernel over   */
/* all sources, self excluded by index at :119-124).                         */
/* ------------------------------------------------------------------------- */
void or_direct(int64_t n, const double *x, const double *y, const double *z,
               const double *q, int64_t ntest, const int64_t *testidx,
               double eps2, double sign, int nthreads,
               double *ex, double *ey, double *ez, double *pot)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t j = 0; j < ntest; ++j) {
        int64_t tj = testidx[j];
        double e0 = 0.0, e1 = 0.0, e2 = 0.0, ph = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            if (i == tj) continue;
            double d[3] = {x[tj] - x[i], y[tj] - y[i], z[tj] - z[i]};
            double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double exyz[3], phi;
            or_p2p(&q[i], d, d2 + eps2, exyz, &phi);
            e0 += exyz[0]; e1 += exyz[1]; e2 += exyz[2]; ph += phi;
        }
        ex[j] = sign * e0; ey[j] = sign * e1; ez[j] = sign * e2; pot[j] = sign * ph;
    }
}

/* ------------------------------------------------------------------------- */
/* Frozen interaction lists and the adjoint of the walk (no reference         */
/* counterpart: the reference has no derivatives).  A list holds, per tile,  */
/* the nodes the group walk interacts with, in traversal order: a leaf is a  */
/* P2P entry, an accepted twig an M2P entry.  Holding the lists and the tree */
/* topology fixed, the fields are smooth functions of the particle positions */
/* and charges; or_eval_lists evaluates them (moments recomputed from the    */
/* particles), or_adjoint_lists returns their exact gradient.                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    const double *x, *y, *z;
    const int32_t *nlevel; const int64_t *nfc, *nns; const double *mom;
    const double *b2; double theta2;
    int64_t t0, np, cnt, cap; int64_t *out;
} wl_ctx;

/* Same decisions as walk_aux (module_tree_walk_simple.f90:301-384). */
static void lists_aux(wl_ctx *w, int64_t node)
{
    if (w->nfc[node] < 0) {
        if (w->out && w->cnt < w->cap) w->out[w->cnt] = node;
        w->cnt++;
        return;
    }
    const double *m = w->mom + OR_MOM * node;
    for (int64_t ip = 0; ip < w->np; ++ip) {
        int64_t p = w->t0 + ip;
        double d0 = w->x[p] - m[OR_COCX], d1 = w->y[p] - m[OR_COCY], d2 = w->z[p] - m[OR_COCZ];
        double dd = d0 * d0 + d1 * d1 + d2 * d2;
        if (!or_mac(w->theta2, dd, w->b2[w->nlevel[node]], m)) {
            for (int64_t c = w->nfc[node]; c >= 0; c = w->nns[c]) lists_aux(w, c);
            return;
        }
    }
    if (w->out && w->cnt < w->cap) w->out[w->cnt] = node;
    w->cnt++;
}

int64_t or_walk_lists(const double *x, const double *y, const double *z,
                      const int32_t *nlevel, const int64_t *nfirst_child,
                      const int64_t *nnext_sibling, const double *nmom,
                      const double b2[OR_MAXLEVEL + 1], double theta2,
                      int64_t ntiles, const int64_t *tile_first, const int64_t *tile_count,
                      int64_t cap, int64_t *list_off, int64_t *list_node)
{
    wl_ctx w = {x, y, z, nlevel, nfirst_child, nnext_sibling, nmom, b2, theta2, 0, 0, 0, cap,
                list_node};
    for (int64_t t = 0; t < ntiles; ++t) {
        if (list_off) list_off[t] = w.cnt;
        w.t0 = tile_first[t];
        w.np = tile_count[t];
        lists_aux(&w, 0);
    }
    if (list_off) list_off[ntiles] = w.cnt;
    return w.cnt > cap ? -w.cnt : w.cnt;
}

/* Moments of every node for a fixed topology (pre-order numbering: a child
 * has a larger index than its parent), leaves from their particle
 * (coulomb/module_interaction_specific.f90:67-79), twigs by
 * shift_multipoles_up. */
void or_tree_moments(int64_t nnodes, const int64_t *nfirst, const int64_t *nfirst_child,
                     const int64_t *nnext_sibling, const double *x, const double *y,
                     const double *z, const double *q, double *nmom)
{
    for (int64_t i = nnodes - 1; i >= 0; --i) {
        double *m = nmom + OR_MOM * i;
        if (nfirst_child[i] < 0) {
            int64_t p = nfirst[i];
            memset(m, 0, sizeof(double) * OR_MOM);
            m[OR_COCX] = x[p]; m[OR_COCY] = y[p]; m[OR_COCZ] = z[p];
            m[OR_CHARGE] = q[p]; m[OR_ABSCHARGE] = fabs(q[p]);
            continue;
        }
        const double *cm[8];
        int nch = 0;
        for (int64_t c = nfirst_child[i]; c >= 0 && nch < 8; c = nnext_sibling[c]) cm[nch++] = nmom + OR_MOM * c;
        shift_multipoles_up(m, cm, nch);
    }
}

/* Fields from frozen lists.  A P2P entry skips the target's own leaf (by
 * index; the walk skips d^2 = 0, the same for distinct positions). */
void or_eval_lists(const double *x, const double *y, const double *z,
                   const int64_t *nfirst, const int64_t *nfirst_child, const double *nmom,
                   double eps2, double sign, int64_t ntiles, const int64_t *tile_first,
                   const int64_t *tile_count, const int64_t *list_off, const int64_t *list_node,
                   double *ex, double *ey, double *ez, double *pot)
{
    for (int64_t t = 0; t < ntiles; ++t) {
        for (int64_t ip = 0; ip < tile_count[t]; ++ip) {
            int64_t p = tile_first[t] + ip;
            double e[3] = {0.0, 0.0, 0.0}, ph = 0.0;
            for (int64_t k = list_off[t]; k < list_off[t + 1]; ++k) {
                const int64_t node = list_node[k];
                const double *m = nmom + OR_MOM * node;
                double d[3] = {x[p] - m[OR_COCX], y[p] - m[OR_COCY], z[p] - m[OR_COCZ]};
                double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                double exyz[3], phi;
                if (nfirst_child[node] < 0) {
                    if (nfirst[node] == p) continue;
                    or_p2p(&m[OR_CHARGE], d, d2 + eps2, exyz, &phi);
                } else {
                    or_m2p(m, d, d2 + eps2, exyz, &phi);
                }
                e[0] += sign * exyz[0]; e[1] += sign * exyz[1]; e[2] += sign * exyz[2];
                ph += sign * phi;
            }
            ex[p] = e[0]; ey[p] = e[1]; ez[p] = e[2]; pot[p] = ph;
        }
    }
}

/* Adjoint of one interaction.  With G = (d^2 + eps^2)^(-1/2) the M2P term is
 * phi = q G + p.grad G + Qt : grad grad G and E = -grad phi (the polynomial
 * of module_coulomb_kernels.f90:60-145; Qt = [[qx, xy/2, zx/2], [xy/2, qy,
 * yz/2], [zx/2, yz/2, qz]]).  For the loss term gphi*phi + u.E it returns
 *   gq = dL/dq, gp = dL/dp, M = dL/dQt (symmetric, full double sum),
 *   D  = dL/dd (d = target - centre).
 * P2P is the case p = Qt = 0. */
void or_adjoint_term(const double d[3], double r2, double gphi, const double u[3], double q,
                     const double p[3], const double qt[6], double D[3], double *gq,
                     double gp[3], double M[6])
{
    const double r = sqrt(r2), ri = 1.0 / r;
    const double r3 = ri * ri * ri, r5 = r3 * ri * ri, r7 = r5 * ri * ri, r9 = r7 * ri * ri;
    const double ud = u[0] * d[0] + u[1] * d[1] + u[2] * d[2];
    const double pd = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
    const double up = u[0] * p[0] + u[1] * p[1] + u[2] * p[2];
    /* Qt as a full symmetric matrix */
    const double Q[3][3] = {{qt[0], 0.5 * qt[3], 0.5 * qt[5]},
                            {0.5 * qt[3], qt[1], 0.5 * qt[4]},
                            {0.5 * qt[5], 0.5 * qt[4], qt[2]}};
    const double tr = qt[0] + qt[1] + qt[2];
    double Qd[3], Qu[3];
    for (int k = 0; k < 3; ++k) {
        Qd[k] = Q[k][0] * d[0] + Q[k][1] * d[1] + Q[k][2] * d[2];
        Qu[k] = Q[k][0] * u[0] + Q[k][1] * u[1] + Q[k][2] * u[2];
    }
    const double dQd = d[0] * Qd[0] + d[1] * Qd[1] + d[2] * Qd[2];
    const double uQd = u[0] * Qd[0] + u[1] * Qd[1] + u[2] * Qd[2];
    *gq = gphi * ri + ud * r3;
    for (int k = 0; k < 3; ++k) gp[k] = (u[k] - gphi * d[k]) * r3 - 3.0 * ud * d[k] * r5;
    static const int mi[6] = {0, 1, 2, 0, 1, 2}, ni[6] = {0, 1, 2, 1, 2, 0};
    for (int k = 0; k < 6; ++k) {
        const int a = mi[k], b = ni[k];
        const double dab = a == b ? 1.0 : 0.0;
        M[k] = gphi * (-dab * r3 + 3.0 * d[a] * d[b] * r5)
               - (3.0 * r5 * (u[a] * d[b] + u[b] * d[a] + dab * ud) - 15.0 * r7 * ud * d[a] * d[b]);
    }
    for (int k = 0; k < 3; ++k) {
        double v = q * gp[k];
        v += gphi * (-p[k] * r3 + 3.0 * d[k] * pd * r5)
             - (3.0 * r5 * (u[k] * pd + p[k] * ud + d[k] * up) - 15.0 * r7 * d[k] * ud * pd);
        v += gphi * (3.0 * r5 * (2.0 * Qd[k] + tr * d[k]) - 15.0 * r7 * d[k] * dQd)
             - (3.0 * r5 * (u[k] * tr + 2.0 * Qu[k])
                - 15.0 * r7 * (u[k] * dQd + 2.0 * ud * Qd[k] + 2.0 * d[k] * uQd + d[k] * ud * tr)
                + 105.0 * r9 * d[k] * ud * dQd);
        D[k] = v;
    }
}

/* Gradient of L = sum_i gpot_i pot_i + gE_i . E_i (the or_eval_lists
 * outputs, sign included) with respect to the sorted particles' x, y, z, q.
 * Node moments enter through their closed forms over the node's particles
 * (q = sum q_j, p = sum q_j s_j, Qt = 1/2 sum q_j s_j s_j^T, s_j = c - x_j,
 * c = sum |q_j| x_j / sum |q_j|); a node whose charges are all zero keeps its
 * centre fixed. */
void or_adjoint_lists(int64_t n, const double *x, const double *y, const double *z,
                      const double *q, int64_t nnodes, const int64_t *nfirst,
                      const int64_t *nleaves, const int64_t *nfirst_child, const double *nmom,
                      double eps2, double sign, int64_t ntiles, const int64_t *tile_first,
                      const int64_t *tile_count, const int64_t *list_off,
                      const int64_t *list_node, const double *gex, const double *gey,
                      const double *gez, const double *gpot, double *gx, double *gy, double *gz,
                      double *gq)
{
    double *adj = (double *)calloc((size_t)nnodes * 16, sizeof(double));
    for (int64_t i = 0; i < n; ++i) gx[i] = gy[i] = gz[i] = gq[i] = 0.0;
    static const double zero3[3] = {0.0, 0.0, 0.0}, zero6[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t t = 0; t < ntiles; ++t) {
        for (int64_t ip = 0; ip < tile_count[t]; ++ip) {
            const int64_t p = tile_first[t] + ip;
            const double gphi = sign * gpot[p];
            const double u[3] = {sign * gex[p], sign * gey[p], sign * gez[p]};
            for (int64_t k = list_off[t]; k < list_off[t + 1]; ++k) {
                const int64_t node = list_node[k];
                const double *m = nmom + OR_MOM * node;
                double d[3] = {x[p] - m[OR_COCX], y[p] - m[OR_COCY], z[p] - m[OR_COCZ]};
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps2;
                double D[3], g, gp[3], M[6];
                if (nfirst_child[node] < 0) {
                    const int64_t s = nfirst[node];
                    if (s == p) continue;
                    or_adjoint_term(d, r2, gphi, u, q[s], zero3, zero6, D, &g, gp, M);
                    gx[s] -= D[0]; gy[s] -= D[1]; gz[s] -= D[2];
                    gq[s] += g;
                } else {
                    const double pm[3] = {m[OR_DIPX], m[OR_DIPY], m[OR_DIPZ]};
                    const double qt[6] = {m[OR_QUADX], m[OR_QUADY], m[OR_QUADZ],
                                          m[OR_XYQUAD], m[OR_YZQUAD], m[OR_ZXQUAD]};
                    or_adjoint_term(d, r2, gphi, u, m[OR_CHARGE], pm, qt, D, &g, gp, M);
                    double *a = adj + 16 * node;
                    a[0] += g;
                    for (int j = 0; j < 3; ++j) a[1 + j] += gp[j];
                    for (int j = 0; j < 6; ++j) a[4 + j] += M[j];
                    for (int j = 0; j < 3; ++j) a[10 + j] += D[j];
                }
                gx[p] += D[0]; gy[p] += D[1]; gz[p] += D[2];
            }
        }
    }
    /* node centre adjoint: dL/dc = q gp + M p - sum D */
    for (int64_t i = 0; i < nnodes; ++i) {
        if (nfirst_child[i] < 0) continue;
        double *a = adj + 16 * i;
        const double *m = nmom + OR_MOM * i;
        const double Mf[3][3] = {{a[4], a[7], a[9]}, {a[7], a[5], a[8]}, {a[9], a[8], a[6]}};
        for (int k = 0; k < 3; ++k) {
            double mp = Mf[k][0] * m[OR_DIPX] + Mf[k][1] * m[OR_DIPY] + Mf[k][2] * m[OR_DIPZ];
            a[13 + k] = m[OR_CHARGE] * a[1 + k] + mp - a[10 + k];
        }
    }
    /* every twig to its particles */
    for (int64_t i = 0; i < nnodes; ++i) {
        if (nfirst_child[i] < 0) continue;
        const double *a = adj + 16 * i;
        const double *m = nmom + OR_MOM * i;
        const double A = m[OR_ABSCHARGE];
        const double Mf[3][3] = {{a[4], a[7], a[9]}, {a[7], a[5], a[8]}, {a[9], a[8], a[6]}};
        for (int64_t j = nfirst[i]; j < nfirst[i] + nleaves[i]; ++j) {
            const double xj[3] = {x[j], y[j], z[j]};
            double s[3], Ms[3];
            for (int k = 0; k < 3; ++k) s[k] = m[OR_COCX + k] - xj[k];
            for (int k = 0; k < 3; ++k) Ms[k] = Mf[k][0] * s[0] + Mf[k][1] * s[1] + Mf[k][2] * s[2];
            double dq = a[0] + a[1] * s[0] + a[2] * s[1] + a[3] * s[2]
                        + 0.5 * (s[0] * Ms[0] + s[1] * Ms[1] + s[2] * Ms[2]);
            double dx[3];
            for (int k = 0; k < 3; ++k) dx[k] = -q[j] * (a[1 + k] + Ms[k]);
            if (A != 0.0) {
                const double sg = q[j] > 0.0 ? 1.0 : (q[j] < 0.0 ? -1.0 : 0.0);
                dq += sg / A * (-(s[0] * a[13]) - s[1] * a[14] - s[2] * a[15]);
                for (int k = 0; k < 3; ++k) dx[k] += fabs(q[j]) / A * a[13 + k];
            }
            gq[j] += dq;
            gx[j] += dx[0]; gy[j] += dx[1]; gz[j] += dx[2];
        }
    }
    free(adj);
}
