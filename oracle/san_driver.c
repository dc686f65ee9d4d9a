/* san_driver.c -- runs every stage of the oracle (the CPU restatement, test
 * infrastructure) on a small system for an AddressSanitizer +
 * UndefinedBehaviorSanitizer build (make -C oracle san): box, keys, presort
 * with coincident particles (duplicate fix-up), tree, tiles, group walk,
 * interaction lists and their evaluation, adjoint, direct sum, lattice
 * coefficients, mirror-box walk and the lattice far field with every
 * extrinsic correction.  Each stage's result gets a cheap consistency check;
 * the sanitizers report any out-of-bounds access, use after free, leak or
 * undefined operation.  Prints "san ok" and exits 0 when all checks pass. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pepc_oracle.h"

static int fails = 0;
#define CHECK(cond, what)                                     \
    do {                                                      \
        if (!(cond)) {                                        \
            fprintf(stderr, "san_driver: check failed: %s\n", what); \
            fails++;                                          \
        }                                                     \
    } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand(void)   /* splitmix64 -> [0, 1) */
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void *xalloc(size_t bytes)
{
    void *p = calloc(1, bytes ? bytes : 1);
    if (!p) { fprintf(stderr, "san_driver: out of memory\n"); exit(2); }
    return p;
}

int main(void)
{
    const int64_t n = 3000;
    double *x = xalloc(sizeof(double) * n), *y = xalloc(sizeof(double) * n);
    double *z = xalloc(sizeof(double) * n), *q = xalloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = urand(); y[i] = urand(); z[i] = urand();
        q[i] = (i % 2) ? -1.0 : 1.0;
    }
    for (int64_t i = 0; i < 40; ++i) {   /* coincident pairs: equal keys, fix-up */
        x[n - 1 - i] = x[i]; y[n - 1 - i] = y[i]; z[n - 1 - i] = z[i];
    }

    /* box, keys, presort, duplicate fix-up */
    or_box b;
    uint64_t *keys = xalloc(sizeof(uint64_t) * n);
    int64_t *perm = xalloc(sizeof(int64_t) * n), nfixed = 0;
    double *xs = xalloc(sizeof(double) * n), *ys = xalloc(sizeof(double) * n);
    double *zs = xalloc(sizeof(double) * n), *qs = xalloc(sizeof(double) * n);
    CHECK(or_domain_decompose(n, x, y, z, q, 0, &b, keys, perm, xs, ys, zs, qs, &nfixed) == 0,
          "domain_decompose");
    CHECK(nfixed > 0, "coincident particles moved by the fix-up");
    for (int64_t i = 1; i < n; ++i) CHECK(keys[i] > keys[i - 1], "keys strictly increasing");
    {
        int64_t ic[3];
        or_key_to_intcoord_hilbert(keys[0], ic);
        CHECK(or_intcoord_to_key_hilbert3d(ic[0], ic[1], ic[2]) == keys[0], "key round trip");
        CHECK(or_level_from_key(keys[0]) == OR_MAXLEVEL, "particle key level");
    }

    /* tree */
    const int64_t cap = 4 * n + 64;
    uint64_t *nkey = xalloc(sizeof(uint64_t) * cap);
    int32_t *nlevel = xalloc(sizeof(int32_t) * cap), *nnchild = xalloc(sizeof(int32_t) * cap);
    int64_t *nleaves = xalloc(sizeof(int64_t) * cap), *nfirst = xalloc(sizeof(int64_t) * cap);
    int64_t *nfc = xalloc(sizeof(int64_t) * cap), *nns = xalloc(sizeof(int64_t) * cap);
    double *nmom = xalloc(sizeof(double) * cap * OR_MOM);
    int64_t *node_leaf = xalloc(sizeof(int64_t) * n), counts[3] = {0, 0, 0};
    CHECK(or_tree_build(n, keys, xs, ys, zs, qs, cap, nkey, nlevel, nleaves, nfirst, nfc, nns,
                        nnchild, nmom, node_leaf, counts) >= 0, "tree_build");
    const int64_t nnodes = counts[0];
    CHECK(nnodes > n && nleaves[0] == n, "root holds every particle");
    double b2[OR_MAXLEVEL + 1];
    or_boxlength2(&b, b2);

    /* tiles, group walk, direct sum */
    int64_t *tf = xalloc(sizeof(int64_t) * n), *tc = xalloc(sizeof(int64_t) * n);
    const int64_t nt = or_walk_tiles(n, keys, 16, 32, tf, tc);
    int64_t covered = 0;
    for (int64_t t = 0; t < nt; ++t) covered += tc[t];
    CHECK(covered == n, "tiles partition the particles");
    double *ex = xalloc(sizeof(double) * n), *ey = xalloc(sizeof(double) * n);
    double *ez = xalloc(sizeof(double) * n), *pot = xalloc(sizeof(double) * n);
    double *work = xalloc(sizeof(double) * n);
    int32_t *cp = xalloc(sizeof(int32_t) * n), *cm = xalloc(sizeof(int32_t) * n);
    int64_t *tmac = xalloc(sizeof(int64_t) * nt), tot[3] = {0, 0, 0};
    const double theta2 = 0.36, eps2 = 1e-4;
    CHECK(or_walk_group(n, xs, ys, zs, nnodes, nlevel, nleaves, nfc, nns, nmom, b2, theta2, eps2,
                        1.0, nt, tf, tc, 1, 2, ex, ey, ez, pot, work, cp, cm, tmac, tot) == 0,
          "walk_group");
    int64_t test[8] = {0, 1, 7, 100, 999, 1500, 2500, n - 1};
    double dx[8], dy[8], dz[8], dp[8];
    or_direct(n, xs, ys, zs, qs, 8, test, eps2, 1.0, 2, dx, dy, dz, dp);
    for (int k = 0; k < 8; ++k) {
        const int64_t i = test[k];
        const double e = sqrt((ex[i] - dx[k]) * (ex[i] - dx[k]) + (ey[i] - dy[k]) * (ey[i] - dy[k]) +
                              (ez[i] - dz[k]) * (ez[i] - dz[k]));
        CHECK(e <= 5e-2 * sqrt(dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k]), "tree vs direct");
    }

    /* interaction lists, their evaluation, the adjoint */
    int64_t *loff = xalloc(sizeof(int64_t) * (nt + 1));
    const int64_t lcap = 4000000;
    int64_t *lnode = xalloc(sizeof(int64_t) * lcap);
    const int64_t nl = or_walk_lists(xs, ys, zs, nlevel, nfc, nns, nmom, b2, theta2, nt, tf, tc,
                                     lcap, loff, lnode);
    CHECK(nl > 0 && nl <= lcap, "walk_lists");
    double *fx = xalloc(sizeof(double) * n), *fy = xalloc(sizeof(double) * n);
    double *fz = xalloc(sizeof(double) * n), *fp = xalloc(sizeof(double) * n);
    or_eval_lists(xs, ys, zs, nfirst, nfc, nmom, eps2, 1.0, nt, tf, tc, loff, lnode, fx, fy, fz, fp);
    for (int64_t i = 0; i < n; ++i) CHECK(fx[i] == ex[i] && fp[i] == pot[i], "lists = walk");
    double *gx = xalloc(sizeof(double) * n), *gy = xalloc(sizeof(double) * n);
    double *gz = xalloc(sizeof(double) * n), *gq = xalloc(sizeof(double) * n);
    or_adjoint_lists(n, xs, ys, zs, qs, nnodes, nfirst, nleaves, nfc, nmom, eps2, 1.0, nt, tf, tc,
                     loff, lnode, ex, ey, ez, pot, gx, gy, gz, gq);
    for (int64_t i = 0; i < n; ++i) CHECK(isfinite(gx[i]) && isfinite(gq[i]), "adjoint finite");

    /* periodic: coefficients, mirror walk, far field with every correction */
    const double lat[9] = {1.3, 0, 0, 0, 1.3, 0, 0, 0, 1.3}, origin[3] = {0, 0, 0};
    const int32_t per[3] = {1, 1, 1};
    double *ml = xalloc(sizeof(double) * 2 * OR_NTAYLOR);
    CHECK(or_lattice_coefficients(lat, per, 1, 32, ml) == 0, "lattice_coefficients");
    const int nb = or_neighbour_boxes(per, 1, NULL);
    int32_t *boxes = xalloc(sizeof(int32_t) * 3 * nb);
    double *vbox = xalloc(sizeof(double) * 3 * nb);
    or_neighbour_boxes(per, 1, boxes);
    for (int k = 0; k < nb; ++k) or_lattice_vect(lat, boxes + 3 * k, vbox + 3 * k);
    CHECK(nb == 27 && vbox[3 * (nb - 1)] == 0.0, "27 boxes, central last");
    double *px = xalloc(sizeof(double) * n), *py = xalloc(sizeof(double) * n);
    double *pz = xalloc(sizeof(double) * n), *qn = xalloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) {   /* the sorted system scaled into the box */
        px[i] = 1.3 * xs[i] / (b.boxmax[0] + 1e-9); py[i] = 1.3 * ys[i] / (b.boxmax[1] + 1e-9);
        pz[i] = 1.3 * zs[i] / (b.boxmax[2] + 1e-9); qn[i] = qs[i];
    }
    CHECK(or_walk_group_boxes(n, xs, ys, zs, nnodes, nlevel, nleaves, nfc, nns, nmom, b2, theta2,
                              eps2, 1.0, nt, tf, tc, 1, 2, nb, vbox, ex, ey, ez, pot, work, cp,
                              cm, tmac, tot) == 0, "walk_group_boxes");
    double *om = xalloc(sizeof(double) * 2 * OR_NMULT), *mu = xalloc(sizeof(double) * 2 * OR_NTAYLOR);
    double fict[16], dip[3], qt = 0.0;
    for (int corr = OR_EXTRINSIC_NONE; corr <= OR_EXTRINSIC_FICTCHARGE; ++corr) {
        int32_t nf = 0;
        or_lattice_timestep(lat, origin, per, ml, n, px, py, pz, qn, corr, om, mu, fict, &nf, dip, &qt);
        CHECK(nf == (corr == OR_EXTRINSIC_FICTCHARGE ? 4 : 0), "fictitious charge count");
        or_lattice_field(lat, origin, per, 1, mu, corr, fict, nf, dip, qt, n, px, py, pz, ex, ey,
                         ez, pot);
        for (int64_t i = 0; i < n; ++i) CHECK(isfinite(ex[i]) && isfinite(pot[i]), "field finite");
    }

    free(x); free(y); free(z); free(q); free(keys); free(perm); free(xs); free(ys); free(zs);
    free(qs); free(nkey); free(nlevel); free(nnchild); free(nleaves); free(nfirst); free(nfc);
    free(nns); free(nmom); free(node_leaf); free(tf); free(tc); free(ex); free(ey); free(ez);
    free(pot); free(work); free(cp); free(cm); free(tmac); free(loff); free(lnode); free(fx);
    free(fy); free(fz); free(fp); free(gx); free(gy); free(gz); free(gq); free(ml); free(boxes);
    free(vbox); free(px); free(py); free(pz); free(qn); free(om); free(mu);
    if (fails) {
        fprintf(stderr, "san_driver: %d checks failed\n", fails);
        return 1;
    }
    printf("san ok: n=%lld nodes=%lld tiles=%lld list entries=%lld\n", (long long)n,
           (long long)nnodes, (long long)nt, (long long)nl);
    return 0;
}
