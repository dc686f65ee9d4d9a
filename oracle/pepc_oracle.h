/*
 * pepc_oracle.h -- CPU restatement of PEPC's Coulomb Barnes-Hut force solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (pepc_amd/libpepc_gpu.so) and the CPU baseline timed by bench.py.  The
 * product never links, loads or calls it; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg do.
 *
 * Every function restates one routine of the reference (Fortran + C) and names
 * the file:line it follows.  All floating-point code is compiled with
 * -ffp-contract=off so that the operation order written here is the one that
 * executes (no fused multiply-add), which is what makes the FP64 moments and the
 * MAC decisions reproducible bit for bit by the device code.
 *
 * Parity pins (how each function is checked, see tests/test_oracle_*.py):
 *   or_box_create           closed form (min/max exact, margin arithmetic)
 *   or_compute_keys         Hilbert invariants: key->coord round trip, face
 *                           adjacency of consecutive cells, bijection on a
 *                           full 2^(3l) grid, hand-computed level-1 order
 *   or_sort_radix_ip        oracle/_ref (the reference's own sort_radix.c /
 *                           sort_insert.c compiled from /root/reference),
 *                           including the order of equal keys
 *   or_dupfix               hand-worked cases of the reference loop
 *   or_tree_build           brute-force prefix counts, moments = direct sums
 *                           over each node's particles (parallel-axis form)
 *   or_walk_group           theta=0 -> direct sum and N leaves per target;
 *                           G=1 vs per-particle recursion; FP64 direct sum
 *   or_set_mac_select       mac_select=2 (never accept) -> direct sum, N
 *                           leaves per target at any theta; mac_select=1 on a
 *                           hand-built pair + far target: accepted exactly
 *                           when theta^2 d^2 > bmax^2 while Barnes-Hut opens it
 *   or_p2p / or_m2p         closed-form Plummer pair; M2P with zero dip/quad
 *                           equals P2P; finite-difference E = -grad(phi)
 *   or_direct               textbook pair sum (numpy float64)
 *   or_walk_lists           list evaluation equals or_walk_group bit for bit
 *   or_tree_moments         equals or_tree_build's moments at the same input
 *   or_adjoint_term/_lists  central finite differences of or_eval_lists with
 *                           frozen lists (h = 1e-5); at theta = 0 the
 *                           closed-form gradient of the pair sum (numpy)
 *   or_neighbour_boxes      hand-listed order of module_mirror_boxes.f90:119-149
 *   or_walk_group_boxes     one box = or_walk_group bit for bit; theta = 0 over
 *                           27 boxes = 3x3x3 image sum (numpy)
 *   or_lattice_coefficients the reference's pretabulated unit-cube table
 *                           (tests/golden/lattice_unit_cube.json); scaling
 *                           law ML(a) = ML(1) / a^(l+1) for a cube of side a
 *   or_lattice_timestep/    near field + lattice field against an Ewald sum
 *   or_lattice_field        (numpy, tin-foil boundary) for random neutral
 *                           systems; rock-salt Madelung constant
 */
#ifndef PEPC_ORACLE_H
#define PEPC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* module_box.f90:33-37 t_box */
typedef struct {
    double boxmin[3];
    double boxmax[3];
    double boxsize[3];
} or_box;

/* Number of doubles per node moment record:
 * coc[3], charge, abs_charge, dip[3], quad[3], xyquad, yzquad, zxquad, bmax, pad
 * (coulomb/module_interaction_specific_types_XYZQ.f90:45-55 order). */
#define OR_MOM 16
enum {
    OR_COCX = 0, OR_COCY, OR_COCZ, OR_CHARGE, OR_ABSCHARGE,
    OR_DIPX, OR_DIPY, OR_DIPZ, OR_QUADX, OR_QUADY, OR_QUADZ,
    OR_XYQUAD, OR_YZQUAD, OR_ZXQUAD, OR_BMAX
};

#define OR_MAXLEVEL 20

/* ---- box and keys --------------------------------------------------------- */
int  or_box_create(int64_t n, const double *x, const double *y, const double *z,
                   int force_cubic, or_box *b);
void or_compute_keys(int64_t n, const double *x, const double *y, const double *z,
                     const or_box *b, uint64_t *keys);
uint64_t or_intcoord_to_key_hilbert3d(int64_t ix, int64_t iy, int64_t iz);
void or_key_to_intcoord_hilbert(uint64_t key, int64_t ic[3]);
void or_key_to_coord(const or_box *b, uint64_t key, double xyz[3]);
int  or_level_from_key(uint64_t key);

/* ---- sort and decomposition ----------------------------------------------- */
void or_sort_radix_ip(int64_t n, uint64_t *keys, int64_t *idx, double *work);
int64_t or_dupfix(int64_t n, uint64_t *keys, double *x, double *y, double *z,
                  const or_box *b);
int or_domain_decompose(int64_t n, const double *x, const double *y, const double *z,
                        const double *q, int force_cubic,
                        or_box *b, uint64_t *keys_sorted, int64_t *perm,
                        double *xs, double *ys, double *zs, double *qs,
                        int64_t *nfixed);

/* ---- tree ------------------------------------------------------------------ */
int64_t or_tree_build(int64_t n, const uint64_t *keys, const double *x, const double *y,
                      const double *z, const double *q, int64_t cap,
                      uint64_t *nkey, int32_t *nlevel, int64_t *nleaves,
                      int64_t *nfirst, int64_t *nfirst_child, int64_t *nnext_sibling,
                      int32_t *nnchild, double *nmom, int64_t *node_leaf,
                      int64_t *counts_out);
void or_boxlength2(const or_box *b, double b2[OR_MAXLEVEL + 1]);

/* ---- tiles and walks ------------------------------------------------------- */
/* MAC of every walk (coulomb/module_interaction_specific.f90:256-275): 0 =
 * Barnes-Hut (default), 1 = min(bmax^2, 3 boxlength2), other = never accept. */
void or_set_mac_select(int mac_select);
int or_get_mac_select(void);
int64_t or_walk_tiles(int64_t n, const uint64_t *keys, int64_t tile_size, int64_t pack,
                      int64_t *tile_first, int64_t *tile_count);
int or_walk_group(int64_t n, const double *x, const double *y, const double *z,
                  int64_t nnodes, const int32_t *nlevel, const int64_t *nleaves,
                  const int64_t *nfirst_child, const int64_t *nnext_sibling,
                  const double *nmom, const double b2[OR_MAXLEVEL + 1],
                  double theta2, double eps2, double sign,
                  int64_t ntiles, const int64_t *tile_first, const int64_t *tile_count,
                  int64_t tile_stride, int nthreads,
                  double *ex, double *ey, double *ez, double *pot, double *work,
                  int32_t *cnt_p2p, int32_t *cnt_m2p, int64_t *tile_mac,
                  int64_t *totals);

int or_walk_group_boxes(int64_t n, const double *x, const double *y, const double *z,
                        int64_t nnodes, const int32_t *nlevel, const int64_t *nleaves,
                        const int64_t *nfirst_child, const int64_t *nnext_sibling,
                        const double *nmom, const double b2[OR_MAXLEVEL + 1],
                        double theta2, double eps2, double sign,
                        int64_t ntiles, const int64_t *tile_first, const int64_t *tile_count,
                        int64_t tile_stride, int nthreads, int64_t nbox, const double *vbox,
                        double *ex, double *ey, double *ez, double *pot, double *work,
                        int32_t *cnt_p2p, int32_t *cnt_m2p, int64_t *tile_mac,
                        int64_t *totals);

/* ---- frozen lists and the adjoint (no reference counterpart) --------------- */
int64_t or_walk_lists(const double *x, const double *y, const double *z,
                      const int32_t *nlevel, const int64_t *nfirst_child,
                      const int64_t *nnext_sibling, const double *nmom,
                      const double b2[OR_MAXLEVEL + 1], double theta2,
                      int64_t ntiles, const int64_t *tile_first, const int64_t *tile_count,
                      int64_t cap, int64_t *list_off, int64_t *list_node);
void or_tree_moments(int64_t nnodes, const int64_t *nfirst, const int64_t *nfirst_child,
                     const int64_t *nnext_sibling, const double *x, const double *y,
                     const double *z, const double *q, double *nmom);
void or_eval_lists(const double *x, const double *y, const double *z,
                   const int64_t *nfirst, const int64_t *nfirst_child, const double *nmom,
                   double eps2, double sign, int64_t ntiles, const int64_t *tile_first,
                   const int64_t *tile_count, const int64_t *list_off, const int64_t *list_node,
                   double *ex, double *ey, double *ez, double *pot);
void or_adjoint_term(const double d[3], double r2, double gphi, const double u[3], double q,
                     const double p[3], const double qt[6], double D[3], double *gq,
                     double gp[3], double M[6]);
void or_adjoint_lists(int64_t n, const double *x, const double *y, const double *z,
                      const double *q, int64_t nnodes, const int64_t *nfirst,
                      const int64_t *nleaves, const int64_t *nfirst_child, const double *nmom,
                      double eps2, double sign, int64_t ntiles, const int64_t *tile_first,
                      const int64_t *tile_count, const int64_t *list_off,
                      const int64_t *list_node, const double *gex, const double *gey,
                      const double *gez, const double *gpot, double *gx, double *gy, double *gz,
                      double *gq);

/* ---- periodic boundaries (pepc_oracle_periodic.c) -------------------------- */
#define OR_LMAX_MULT 20          /* Lmax_multipole, module_fmm_framework.f90:70 */
#define OR_LMAX_TAYLOR 40        /* Lmax_taylor, :71 */
#define OR_NMULT 231             /* (LM+1)(LM+2)/2 */
#define OR_NTAYLOR 861           /* (LT+1)(LT+2)/2 */
int  or_neighbour_boxes(const int32_t per[3], int32_t layers, int32_t *boxes);
void or_lattice_vect(const double lat[9], const int32_t ijk[3], double v[3]);
int  or_lattice_coefficients(const double lat[9], const int32_t per[3], int32_t layers,
                             int32_t maxiter, double *ml);
/* extrinsic-to-intrinsic corrections, module_fmm_framework.f90:42-46 */
#define OR_EXTRINSIC_NONE 0
#define OR_EXTRINSIC_REDLACK 1
#define OR_EXTRINSIC_FICTCHARGE 2
void or_lattice_timestep(const double lat[9], const double origin[3], const int32_t per[3],
                         const double *ml, int64_t n, const double *x, const double *y,
                         const double *z, const double *q, int32_t correction, double *omega,
                         double *mu, double *fict, int32_t *nfict, double dipole[3],
                         double *quad_trace);
void or_lattice_field(const double lat[9], const double origin[3], const int32_t per[3],
                      int32_t layers, const double *mu, int32_t correction, const double *fict,
                      int32_t nfict, const double dipole[3], double quad_trace, int64_t n,
                      const double *x, const double *y, const double *z, double *ex,
                      double *ey, double *ez, double *pot);

/* ---- kernels and direct sum ------------------------------------------------ */
void or_p2p(const double *t_charge, const double d[3], double dist2,
            double exyz[3], double *phi);
void or_m2p(const double *mom, const double d[3], double dist2,
            double exyz[3], double *phi);
void or_direct(int64_t n, const double *x, const double *y, const double *z,
               const double *q, int64_t ntest, const int64_t *testidx,
               double eps2, double sign, int nthreads,
               double *ex, double *ey, double *ez, double *pot);

#ifdef __cplusplus
}
#endif
#endif
