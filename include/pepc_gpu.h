/*
 * pepc_gpu.h -- C-ABI of the MI355X Coulomb Barnes-Hut force solve.
 *
 * Drop-in replacement for the frontends' call into PEPC's treecode
 * (pepc_grow_tree + pepc_traverse_tree on one rank) for the Coulomb backend
 * with the XYZQ particle type.  Plain C types only: double arrays in
 * structure-of-arrays form, 64-bit sizes, int status returns.  A Fortran
 * frontend binds these with iso_c_binding (pepc_amd/fortran/module_pepc_gpu.f90);
 * Python binds them with ctypes (pepc_amd/_lib.py).
 *
 * Status codes: 0 = success, negative = error (pepc_gpu_last_error() gives
 * the message).  The reference has no status returns at its Fortran API; its
 * fatal errors abort through DEBUG_ERROR (pepc_debug.h:28-35) and its C calls
 * return int checked by ERROR_ON_FAIL (pepc_debug.h:54-61).  The Fortran shim
 * maps a nonzero status to the same kind of abort.
 *
 * Ordering.  Like the reference on one rank (module_domains.f90:204-234),
 * grow sorts the particles by space-filling key; results are returned in that
 * sorted order and perm[i] names the input index of sorted particle i
 * (the reference's indxl(i) - 1).  pepc-mini/pepc-essential keep the sorted
 * order (src/frontends/pepc-mini/pepc.f90:85); pepc_restore_particles undoes it.
 *
 * Threading.  One context is used by one host thread at a time, as the
 * reference API is (MPI_THREAD_SERIALIZED, module_pepc.f90:88).  All work of
 * a context is ordered on its HIP stream.
 */
#ifndef PEPC_GPU_H
#define PEPC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PEPC_GPU_OK = 0,
    PEPC_GPU_EINVAL = -1,     /* bad argument or call order */
    PEPC_GPU_EHIP = -2,       /* HIP runtime error (no device, launch failure) */
    PEPC_GPU_ENOMEM = -3,     /* device allocation failed */
    PEPC_GPU_EOVERFLOW = -4,  /* node pool too small (raise node_factor, cf. np_mult) */
    PEPC_GPU_EKEYS = -5,      /* key space exhausted / unresolvable duplicate keys */
    PEPC_GPU_EBOX = -6,       /* non-finite coordinates or a degenerate box */
    PEPC_GPU_EWALK = -7       /* walk stack overflow */
};

enum { PEPC_GPU_FP32_KAHAN = 0, PEPC_GPU_FP64 = 1 };

/* Solver parameters; the &calc_force_coulomb and &libpepc namelist entries
 * of the reference (coulomb/module_interaction_specific.f90:35-43,
 * module_libpepc_main.f90:44) that apply to this path. */
typedef struct {
    double theta2;          /* BH opening parameter squared (default 0.36) */
    double eps2;            /* Plummer softening squared (default 0) */
    double sign;            /* +1 Coulomb; -1 attractive (gravity-like) */
    double node_factor;     /* node capacity per particle (np_mult analogue, default 2.5);
                               node_factor * (n + n/8) must stay below 2^31 nodes (EINVAL) */
    int32_t mac_select;     /* 0 = Barnes-Hut theta2 d^2 > boxlength2(level); 1 = Bmax,
                               theta2 d^2 > min(bmax^2, 3 boxlength2); any other value
                               never accepts (N^2), as coulomb/module_interaction_specific.f90:256-275 */
    int32_t group_size;     /* walk tile size G, 1..64 (16 = module_tree_walk_simple TILE_SIZE) */
    int32_t precision;      /* PEPC_GPU_FP32_KAHAN or PEPC_GPU_FP64 */
    int32_t force_cubic_domain;  /* module_box.f90:29 */
    int32_t walk_blocks;    /* persistent walk grid; 0 = fill the device */
    int32_t device;         /* HIP device ordinal */
    int32_t group_pack;     /* 0 = reference tiles; 1..64 = merge runs of sibling tiles
                               up to this many particles per group (no reference
                               counterpart; fuller wavefronts) */
    int32_t list_tile;      /* M2P interaction list entries staged in LDS per flush:
                               0 = default (64 FP32, 32 FP64); FP32 64|128, FP64 32|64 */
} pepc_gpu_config;

/* Counters of the last grow/traverse (tree_stats, module_tree.f90:622-699;
 * tree_walk_statistics, module_tree_walk_simple.f90:63-115). */
typedef struct {
    int64_t n;
    int64_t nnodes, nleaf, ntwig, ngroups;
    int64_t dup_fixed;          /* particles moved by the duplicate-key fix-up */
    uint64_t interactions;      /* sum over targets of P2P + M2P (self included) */
    uint64_t mac_evaluations;   /* per-tile MAC tests, first failure ends a node */
    uint64_t p2p, m2p;
    uint64_t amb_nodes;         /* twigs the group bounding box left undecided */
    uint64_t walk_steps;        /* traversal steps of all groups */
    uint32_t flags;
    int32_t maxlevel;
    uint32_t level_offset[22];  /* first node of each level in the device layout */
} pepc_gpu_stats;

/* Fills *cfg with the defaults above. */
void pepc_gpu_config_default(pepc_gpu_config *cfg);

/* Replaces pepc_initialize + pepc_prepare (module_pepc.f90:70-175, 269-288):
 * binds a device and creates the context's stream. */
int pepc_gpu_create(void **ctx, const pepc_gpu_config *cfg);

/* Replaces pepc_finalize (module_pepc.f90:177-200). */
int pepc_gpu_destroy(void *ctx);

/* Copies the last error message (NUL-terminated) into buf. */
int pepc_gpu_last_error(const void *ctx, char *buf, size_t len);

/* Orders the context's work on a caller stream (hipStream_t) instead of its own. */
int pepc_gpu_set_stream(void *ctx, void *hip_stream);

/* Replaces pepc_grow_tree (module_pepc.f90:435-446 -> tree_grow,
 * module_tree_grow.f90:37-138), host arrays.  Inputs x,y,z,q[n] in any
 * order; outputs (each may be NULL): keys_out[n] sorted keys, perm_out[n],
 * xs/ys/zs[n] sorted positions after the duplicate-key fix-up
 * (module_domains.f90:299-305 rewrites moved particles).  Synchronous.
 * When the tree needs more nodes than the pool holds, this host form
 * enlarges the pool to the count the build found and grows again, so it
 * never returns PEPC_GPU_EOVERFLOW (the device forms do, and leave the pool
 * size to node_factor and pepc_gpu_reserve_nodes). */
int pepc_gpu_grow(void *ctx, int64_t n, const double *x, const double *y, const double *z,
                  const double *q, uint64_t *keys_out, int64_t *perm_out,
                  double *xs, double *ys, double *zs);

/* Replaces pepc_traverse_tree (module_pepc.f90:459-468 ->
 * libpepc_traverse_tree, module_libpepc_main.f90:161-200) for the grown
 * particles.  Outputs in sorted order (each may be NULL): field ex,ey,ez,
 * potential pot (particle%results), work = 1 + interactions (particle%work,
 * module_domains.f90:318 + walk increments).  Synchronous. */
int pepc_gpu_traverse(void *ctx, double *ex, double *ey, double *ez, double *pot, double *work);

/* Device-pointer forms of grow and traverse: asynchronous on the context's
 * stream; call pepc_gpu_sync() to wait and check for errors. */
int pepc_gpu_grow_device(void *ctx, int64_t n, const double *x, const double *y,
                         const double *z, const double *q);
int pepc_gpu_traverse_device(void *ctx, double *ex, double *ey, double *ez, double *pot,
                             double *work);

/* Traversal of other particles: replaces pepc_traverse_tree(particles) for
 * a particle array that is not the grown one (pepc_grow_and_traverse_for_others,
 * module_pepc.f90:394-433; libpepc_traverse_tree :161-200).  Fields and
 * potentials at m sink positions x,y,z (device arrays, any order; outputs in
 * the same order) from the grown tree, periodic boundaries included.
 * PEPC_GPU_SINKS_PER_PARTICLE: every sink walks alone, the reference's rule
 * for particles without keys (module_tree_walk_simple.f90:155-159), so
 * counts equal the reference's.  PEPC_GPU_SINKS_GROUPED (no reference
 * counterpart): sinks sorted along the source box's Hilbert curve (clamped
 * into the box) walk in runs of group_size; same physics, fewer traversals,
 * other counts.  A sink at exactly a source's position skips that source
 * (zero separation is self, :331-340).  work[i] = interactions of sink i
 * (the reference adds them to %work).  Not with batched systems.
 * The _device form is asynchronous; the host form copies and synchronises. */
enum { PEPC_GPU_SINKS_PER_PARTICLE = 0, PEPC_GPU_SINKS_GROUPED = 1 };
int pepc_gpu_traverse_sinks_device(void *ctx, int64_t m, const double *x, const double *y,
                                   const double *z, int32_t mode, double *ex, double *ey,
                                   double *ez, double *pot, double *work);
int pepc_gpu_traverse_sinks(void *ctx, int64_t m, const double *x, const double *y,
                            const double *z, int32_t mode, double *ex, double *ey, double *ez,
                            double *pot, double *work);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): pepc_gpu_traverse_sinks_device with the same arguments and
 * results, but its sort buffers, staging and counters come from a static
 * device array of 26.2 MB per device instead of the context's growing sink
 * buffers, so the call allocates nothing.  At most 262144 sinks (more return
 * PEPC_GPU_EINVAL).  All calls on one device share that array: calls that
 * overlap on one device (any context) race.  pepc_gpu_read_sink_counts reads
 * the counts of the last call, as after the _device form. */
int pepc_gpu_traverse_sinks_noalloc(void *ctx, int64_t m, const double *x, const double *y,
                                    const double *z, int32_t mode, double *ex, double *ey,
                                    double *ez, double *pot, double *work);
/* Per-sink P2P and M2P counts of the last sink traversal (sinks' order). */
int pepc_gpu_read_sink_counts(void *ctx, int64_t cap, uint32_t *p2p, uint32_t *m2p);

/* Replaces tree_check (module_tree.f90:438-521) behind pepc_check_sanity
 * (module_pepc.f90:504-520): checks the grown tree on the device (child
 * ranges, levels, parent links, particle counts, one leaf per particle, one
 * root per system) and sets *nbad to the number of violations (0 = sound).
 * Synchronous; allocates a small result buffer. */
int pepc_gpu_check_tree(void *ctx, int64_t *nbad);

/* Waits for the context's work, reads its counters into *st (may be NULL)
 * and reports errors raised by the kernels. */
int pepc_gpu_sync(void *ctx, pepc_gpu_stats *st);

/* Lower bound of the node pool: at least `nodes` entries from the next grow
 * on (the pool is node_factor * n otherwise).  tree_create's table size
 * (module_tree.f90:177-186: max(30 nintmax + 4 N, 10^4) for np_mult > 0,
 * |np_mult| 10^4 otherwise, at least 2^15) maps here; the Fortran shim
 * passes it before each grow.  Enlarging the pool drops the current tree. */
int pepc_gpu_reserve_nodes(void *ctx, int64_t nodes);

/* Replaces pepc_timber_tree (module_pepc.f90:538): releases nothing but
 * marks the tree invalid (device buffers are kept for the next step). */
int pepc_gpu_timber(void *ctx);

/* Device pointers to the sorted particle arrays of the last grow (valid
 * until the next grow): keys[n] (uint64), perm[n] (uint32), x,y,z,q[n]. */
int pepc_gpu_sorted_device(void *ctx, const uint64_t **keys, const uint32_t **perm,
                           const double **x, const double **y, const double **z,
                           const double **q);

/* Replaces pepc_restore_particles (module_pepc.f90:525-533 -> domain_restore,
 * module_domains.f90:372-408) for per-particle device arrays: for each of
 * the count arrays, out[k][perm[i]] = in[k][i], i.e. sorted order back to the
 * grow's input order (e.g. traverse results).  Arrays of the grown size, not
 * in place.  Asynchronous on the context's stream. */
int pepc_gpu_restore_device(void *ctx, int64_t count, const double *const *in,
                            double *const *out);

/* Host copies of the sorted particle arrays of the last grow (each may be
 * NULL): keys[n], perm[n] (input index of sorted particle i), x,y,z,q[n]. */
int pepc_gpu_read_sorted(void *ctx, uint64_t *keys, int64_t *perm, double *x, double *y,
                         double *z, double *q);

/* Copies the tree of the last grow to host arrays of capacity cap nodes:
 * link[4*i] = {first child or particle, #children (0 = leaf), #particles,
 * level}, first[i] = first particle, mom[16*i] = moments in the reference
 * layout {coc(3), charge, abs_charge, dip(3), quad(3), xy, yz, zx, bmax, 0}
 * (coulomb/module_interaction_specific_types_XYZQ.f90:45-55).  Returns the
 * node count in *nnodes. */
int pepc_gpu_read_tree(void *ctx, int64_t cap, uint32_t *link, uint32_t *first, double *mom,
                       int64_t *nnodes);

/* Bounding box of system sys (0 for a single system) of the last grow, as
 * box_create computes it (module_box.f90:48-93, t_box :33-38): boxmin[3],
 * boxmax[3], boxsize[3] (each may be NULL).  The tree writers place node
 * boxes with it (key_to_coord, module_spacefilling.f90:445-471). */
int pepc_gpu_read_box(void *ctx, int64_t sys, double *boxmin, double *boxmax, double *boxsize);

/* Copies the walk groups and per-target counts of the last traverse:
 * gfirst/gcount/gmac[ngroups] and p2p/m2p[n] (sorted order). */
int pepc_gpu_read_walk(void *ctx, int64_t gcap, uint32_t *gfirst, uint32_t *gcount,
                       uint32_t *gmac, uint32_t *p2p, uint32_t *m2p, int64_t *ngroups);

/* FP64 direct sum (src/utils/module_directsum.f90:34-172) on device arrays:
 * fields at the ntest targets testidx[] from all n sources, self excluded by
 * index.  With periodic boundaries set on the context, directforce's
 * periodic form: every source in every mirror box, self only in the central
 * box (:80-93), plus the lattice far field at the targets from all sources
 * (:130-137), after which pepc_gpu_lattice_get reports these sources'
 * moments.  Asynchronous on the context's stream. */
int pepc_gpu_direct_device(void *ctx, int64_t n, const double *x, const double *y,
                           const double *z, const double *q, int64_t ntest,
                           const int64_t *testidx, double eps2, double sign, double *ex,
                           double *ey, double *ez, double *pot);

/* Host form of pepc_gpu_direct_device (directforce, src/utils/
 * module_directsum.f90:34-172, open or periodic): host arrays in and
 * out, indices checked against n (EINVAL), synchronous; device memory for
 * the sources is taken for the call and released after it. */
int pepc_gpu_direct(void *ctx, int64_t n, const double *x, const double *y, const double *z,
                    const double *q, int64_t ntest, const int64_t *testidx, double eps2,
                    double sign, double *ex, double *ey, double *ez, double *pot);

/* Device timing of the last grow and traverse in milliseconds (HIP events
 * on the context's stream, recorded by the *_device and host forms). */
int pepc_gpu_timings(void *ctx, double *t_grow_ms, double *t_walk_ms);

/* ---------------------------------------------------------------------------
 * Batched systems (no reference counterpart: the reference solves one system
 * per MPI communicator).  nsys independent systems in one grow/traverse:
 * system s holds particles [offsets[s], offsets[s+1]) of the device arrays
 * (host offsets, nsys + 1 entries, offsets[0] = 0, offsets[nsys] = n,
 * 1 <= nsys <= 65536).  Each system gets its own box, keys, tree and walk,
 * exactly as a separate pepc_gpu_grow_device of that range would; sorted
 * results keep the ranges (system s sorted by key in [offsets[s],
 * offsets[s+1])).  Traverse with pepc_gpu_traverse_device. */
int pepc_gpu_grow_batch_device(void *ctx, int64_t nsys, const int64_t *offsets, const double *x,
                               const double *y, const double *z, const double *q);

/* ---------------------------------------------------------------------------
 * On-device timestep around the solve (the frontends' particle push,
 * pepc-mini module_helper.f90:205-222, kept in HBM).  One step is
 * kick-drift-kick leapfrog: v += dt/2 a; x += dt v; grow + traverse;
 * a = q E / m; v += dt/2 a.  The reference frontend pushes kick-drift
 * (v += dt a; x += dt v) instead.  After every step the particles are in key
 * order, as after pepc_grow_tree; label[i] is the input index of particle i.
 *
 * state_set loads n particles (host arrays, nsys systems as in
 * grow_batch_device; offsets may be NULL when nsys = 1), runs the initial
 * solve and records energies (record 0).  Energies are kept for up to
 * max_records solves: record r, system s = {1/2 sum q phi, 1/2 sum m v^2}. */
int pepc_gpu_state_set(void *ctx, int64_t n, int64_t nsys, const int64_t *offsets,
                       const double *x, const double *y, const double *z, const double *vx,
                       const double *vy, const double *vz, const double *q, const double *m,
                       int64_t max_records);

enum { PEPC_GPU_RUN_GRAPH = 1 };

/* Runs nsteps steps of size dt, asynchronously on the context's stream.  With
 * PEPC_GPU_RUN_GRAPH the step is captured once into a HIP graph (re-captured
 * when dt changes) and replayed; results are identical to the eager form. */
int pepc_gpu_run(void *ctx, int64_t nsteps, double dt, int flags);

/* ---------------------------------------------------------------------------
 * Adjoint of the solve (no reference counterpart).  For a loss L with
 * gex,gey,gez = dL/dE and gpot = dL/dpot at the grown particles (device
 * arrays, sorted order), writes gx,gy,gz = dL/dx and gq = dL/dq (sorted
 * order), FP64, with the tree and the interaction lists of the last grow held
 * fixed (the lists are piecewise constant in the positions).  Particles moved
 * by the duplicate-key fix-up are treated as fixed inputs.  The sums are
 * FP64 atomic additions from concurrent wavefronts, so two calls on the same
 * inputs agree to rounding, not bit for bit.  Asynchronous on the context's
 * stream. */
int pepc_gpu_adjoint_device(void *ctx, const double *gex, const double *gey, const double *gez,
                            const double *gpot, double *gx, double *gy, double *gz,
                            double *gq);

/* Frozen interaction lists.  freeze records, for the last grow, every walk
 * group's interaction list (the nodes of its leaf and accepted-twig
 * interactions, in traversal order, by the reference acceptance rule);
 * *nentries = total entries.  Synchronous.  replay_device then evaluates the
 * frozen lists for new particle data x,y,z,q (device arrays, the grow's
 * INPUT order): gathered through the grow permutation, moments recomputed on
 * the same tree topology (FP64 upward pass), fields and potentials in FP64,
 * sorted order out.  This is the function whose derivative
 * pepc_gpu_adjoint_device returns (central differences of replay check it).
 * Replay replaces the context's particle data; the next grow invalidates the
 * lists. */
int pepc_gpu_freeze(void *ctx, int64_t *nentries);
int pepc_gpu_replay_device(void *ctx, const double *x, const double *y, const double *z,
                           const double *q, double *ex, double *ey, double *ez, double *pot);

/* ---------------------------------------------------------------------------
 * Autotuner (no reference counterpart).  Times grow + traverse of the given
 * device arrays (one system, or nsys batched systems with host offsets) for
 * every candidate of group size {16, 32, 64} x group pack {0, 32, 64} x list tile
 * (FP32 {64, 128}, FP64 {32, 64}) x persistent grid {full, half}: one
 * warm-up and reps timed runs each, device time (HIP events), best of reps.
 * The fastest becomes the context's configuration (read it back with
 * pepc_gpu_get_config); the tree is invalidated, so grow again.  out[0] is
 * the winner, out[1..] every candidate in the order tried (up to cap
 * entries); *ncand = number of candidates.  Synchronous. */
typedef struct {
    int32_t group_size, group_pack, list_tile, walk_blocks;
    double t_grow_ms, t_walk_ms;
} pepc_gpu_tune_entry;

int pepc_gpu_autotune(void *ctx, int64_t n, int64_t nsys, const int64_t *offsets,
                      const double *x, const double *y, const double *z, const double *q,
                      int32_t reps, pepc_gpu_tune_entry *out, int64_t cap, int64_t *ncand);

/* Copies the context's current configuration (walk_blocks = the grid in use). */
int pepc_gpu_get_config(const void *ctx, pepc_gpu_config *cfg);

/* Host copies of the state (each may be NULL), current order. */
int pepc_gpu_state_get(void *ctx, double *x, double *y, double *z, double *vx, double *vy,
                       double *vz, double *q, double *m, int64_t *label);

/* Device pointers to the state: arrays[11] = x y z vx vy vz q m ax ay az,
 * *label = int64 labels.  Valid until the next state_set. */
int pepc_gpu_state_device(void *ctx, double **arrays, const int64_t **label);

/* Waits, then copies min(records, cap) energy records (cap x nsys x 2
 * doubles) to out; *nrec = solves so far (records beyond max_records are
 * not kept). */
int pepc_gpu_read_energies(void *ctx, int64_t cap, double *out, int64_t *nrec);

/* ---------------------------------------------------------------------------
 * Periodic boundaries.  The reference's settings: periodicity and
 * mirror_box_layers of the &libpepc namelist (module_libpepc_main.f90:44),
 * the lattice t_lattice_1..3 and LatticeOrigin (module_mirror_boxes.f90:
 * 30-36; neighbour_boxes_prepare :97-153), include_far_field_if_periodic of
 * &calc_force_coulomb (coulomb/module_interaction_specific.f90:43).
 *
 * With any periodic axis, traverse walks the tree once per neighbour box
 * (lattice vectors within mirror_box_layers shells of the periodic axes, the
 * reference's box order; module_libpepc_main.f90:184-187) and, with
 * far_field, adds the lattice contribution of everything beyond
 * (fmm_framework_timestep + fmm_sum_lattice_force, module_fmm_framework.f90:
 * 132-149, 766-833: lattice coefficients to order 40, central-box multipoles
 * to order 20, well separation ws = mirror_box_layers, the extrinsic-to-
 * intrinsic correction fmm_extrinsic_correction, :42-49, 807-829).  A
 * unit-cube lattice with one mirror layer, periodic on all three axes, uses
 * the reference's pretabulated coefficients (:114-121, :260-508); other
 * lattices, layer counts and periodicities compute them
 * (calc_lattice_coefficients, :154-213, over the periodic axes' neighbour
 * boxes) on the device in set_periodic.  With one or two periodic axes the
 * far field still runs, as the reference's does (do_periodic .and.
 * include_far_field_if_periodic).
 * The on-device timestep wraps positions back into the box after each drift
 * (constrain_periodic, module_mirror_boxes.f90:174-210).
 *
 * Limits: mirror_box_layers 1..3.
 * Not with batched systems, the adjoint or frozen lists (EINVAL).  Particles
 * must lie in the central box (the reference warns, check_lattice_boundaries
 * :261-281). */
/* fmm_extrinsic_correction values (module_fmm_framework.f90:42-46; the
 * reference's MEASUREMENT (3) is unimplemented there and EINVAL here). */
enum {
    PEPC_GPU_EXTRINSIC_NONE = 0,        /* none */
    PEPC_GPU_EXTRINSIC_REDLACK = 1,     /* Redlack-Grindlay dipole and trace terms */
    PEPC_GPU_EXTRINSIC_FICTCHARGE = 2   /* fictitious charges (Kudin), the default */
};
typedef struct {
    int32_t periodicity[3];       /* nonzero = periodic along t_lattice_k */
    int32_t mirror_box_layers;    /* default 1 */
    double t_lattice[9];          /* rows: t_lattice_1, t_lattice_2, t_lattice_3 */
    double origin[3];             /* LatticeOrigin */
    int32_t far_field;            /* include_far_field_if_periodic (default 1) */
    int32_t extrinsic_correction; /* fmm_extrinsic_correction (default FICTCHARGE) */
} pepc_gpu_periodic;

/* Fills *p with the reference defaults: unit cube at the origin, no axis
 * periodic, one layer, far field on, fictitious-charge correction. */
void pepc_gpu_periodic_default(pepc_gpu_periodic *p);

/* Sets (p != NULL) or clears (p == NULL, or no periodic axis) periodic
 * boundaries for the following traverses.  Synchronous. */
int pepc_gpu_set_periodic(void *ctx, const pepc_gpu_periodic *p);

/* Host copies of the far-field state (each may be NULL): ml[2*861] lattice
 * coefficients (interleaved re, im; l <= 40 at l(l+1)/2 + m), and after a
 * traverse mu[2*231] the lattice's local expansion at the box centre,
 * fict[16] the fictitious charges {x, y, z, q} x 4 and *nfict their count. */
int pepc_gpu_lattice_get(void *ctx, double *ml, double *mu, double *fict, int32_t *nfict);

/* constrain_periodic (module_mirror_boxes.f90:174-210) on device arrays:
 * positions of the periodic axes wrapped into the central box.
 * Asynchronous on the context's stream. */
int pepc_gpu_constrain_periodic_device(void *ctx, int64_t n, double *x, double *y, double *z);

#ifdef __cplusplus
}
#endif
#endif
