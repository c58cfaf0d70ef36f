/*
 * treemi.h -- C-ABI of the MI355X Barnes-Hut engine (libtreemi.so).
 *
 * The engine holds one whole plasma domain (electrons + singly charged ions)
 * resident on one GPU behind an opaque handle, and replaces the reference's
 * single-domain force / time-step path:
 *
 *   tb_build          <- BuildLocalTree     tree.f90:39-100 (+ Insert_Tree_e/_i
 *                                            tree.f90:103-447, PSeudoE/I 450-659)
 *   tb_force          <- LocalCoulombSum    tree.f90:662-717 (+ Search_Tree_* 721-1133)
 *   tb_vv1            <- VVerletNotemp1     vverlet.f90:9-44
 *   tb_vv2            <- VVerletNotemp2     vverlet.f90:49-112
 *   tb_set_particles  <- PTEgathering2 / PTIgathering2, main.f90:430-502
 *                        (positions -> device; the MPI gather collapses)
 *   tb_get_particles  <- host view of qel/qion (TYPE PT) for TEMPORAL / Restart
 *   tb_step           <- the time loop body main.f90:126-133 (VV1, Force, VV2)
 *
 * Units are the reference's (data.f90:29-40): um, eV, amu, t0 = 0.1018 ns.
 * Every call returns TB_OK (0) or a TB_ERR_* code; tb_last_error() gives the
 * message.  The reference STOPs on errors; its Fortran binding does the same
 * on a nonzero status.
 *
 * Threading: one host thread per handle; calls on one handle are not
 * re-entrant.  All device work of a handle runs on the handle's own HIP
 * stream; calls that return host data synchronise that stream.
 */
#ifndef TREEMI_H
#define TREEMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TB_OK 0
#define TB_ERR_ARG 1        /* bad argument (null pointer, n < 1, theta <= 0, ...) */
#define TB_ERR_HIP 2        /* HIP runtime error */
#define TB_ERR_DEPTH 3      /* two particles share all 21 key levels (coincident) */
#define TB_ERR_NOMEM 4      /* device allocation failed or over the handle's cap */
#define TB_ERR_STATE 5      /* call order (e.g. tb_force before tb_build) */
#define TB_ERR_NONFINITE 6  /* NaN / Inf position */
#define TB_ERR_NODEV 7      /* no usable GPU */

/* TYPE PT, data.f90:47-50: xx(3), xv(3), ff(3) REAL*8, id INTEGER (+pad) */
typedef struct tb_pt {
    double xx[3];
    double xv[3];
    double ff[3];
    int32_t id;
    int32_t pad_;
} tb_pt;

/* walk variants */
#define TB_WALK_REF64 0   /* per-target FP64 walk, reference MAC / moments / order */
#define TB_WALK_MIXED 1   /* wavefront-group walk, FP32 P2P/M2P, compensated sums */
/* opening criterion for TB_WALK_MIXED */
#define TB_MAC_FIXED 0    /* fixed theta */
#define TB_MAC_LEARNED 1  /* per-interaction theta from the trained MLP (tb_mlp_*) */

/* TYPE PM, data.f90:90-93, plus engine options */
typedef struct tb_config {
    double xm[2];      /* masses, amu: [0] electron, [1] ion          (PM%xm)      */
    double rs;         /* Kelbg length, um                           (PM%rs)      */
    double theta;      /* opening angle of TB_WALK_REF64             (PM%sd_ratio) */
    double theta_mixed;/* opening angle of TB_WALK_MIXED (0: = theta)              */
    double theta_lo;   /* TB_MAC_LEARNED: smallest theta the MLP may pick          */
    double err_target; /* TB_MAC_LEARNED: accepted predicted per-interaction error */
    int32_t walk;      /* TB_WALK_*                                               */
    int32_t mac;       /* TB_MAC_*                                                */
    int32_t group;     /* TB_WALK_MIXED targets per group: 0 = autotuned, else 64, 128 or 256 */
    int32_t device;    /* HIP device ordinal                                       */
    int64_t mem_cap;   /* device bytes the handle may hold (0 = unlimited)         */
    double far_ratio;  /* TB_WALK_MIXED: an accepted cell farther than far_ratio    */
                       /* times the target group's radius from the group's centre */
                       /* goes through the group's local expansion instead of     */
                       /* being evaluated at every target (0: default 3.5, < 0: */
                       /* off, else >= 2)                                        */
    int32_t far_levels;/* TB_WALK_MIXED with the far path: 0 or 1 (default) the   */
                       /* group level only; 2: two levels, a segment of up to 8   */
                       /* neighbouring groups walks the top of the tree once, and */
                       /* cells far from all of it share the segment's local      */
                       /* expansion (same acceptance, fewer visits and M2L pairs; */
                       /* no faster at 4M, sigma 600 um, theta_mixed 0.27)        */
    float cost_weight; /* TB_MAC_LEARNED: weight w of the work a band cell's       */
                       /* acceptance saves.  The cell is accepted when the        */
                       /* predicted log10 error is below log10(err_target) +      */
                       /* w log10(S), S = (estimated cost of opening it - one      */
                       /* M2P) / one M2P, from its children's records (leaves and */
                       /* small cells: P2P per particle; larger cells: one M2P    */
                       /* each); S <= 0 opens it.  0: error only                  */
} tb_config;

typedef struct tb_handle tb_handle;

int tb_create(tb_handle **h, const tb_config *cfg);
int tb_destroy(tb_handle *h);
int tb_set_config(tb_handle *h, const tb_config *cfg);

/* positions, velocities in the reference's index order; forces are not read
 * (the device forces start at zero until the next tb_force) */
int tb_set_particles(tb_handle *h, int32_t ne, const tb_pt *e, int32_t ni, const tb_pt *i);
/* x, v, f back in the order given to tb_set_particles */
int tb_get_particles(tb_handle *h, tb_pt *e, tb_pt *i);

int tb_build(tb_handle *h);
int tb_force(tb_handle *h, double *epot_e, double *epot_i);
int tb_vv1(tb_handle *h, double dt);
/* sums (if sample != 0): [0] mv2el [1] mv2ion [2] Te_rad [3] Ti_rad (vverlet.f90:81-106) */
int tb_vv2(tb_handle *h, double dt, int32_t sample, double *sums);
/*
 * nsteps x (VV1, build, force, VV2); the kick of VV2 and the next VV1 run as
 * one fused kernel.  out (optional, 6 doubles) after the last step:
 * Epot_e, Epot_i, mv2el, mv2ion, Te_rad, Ti_rad.
 */
int tb_step(tb_handle *h, double dt, int32_t nsteps, double *out);

/*
 * Node table of one species' tree (0 electron, 1 ion) in preorder with
 * children in the reference's order 1..8; next[] = preorder index after the
 * subtree; pt[] = particle index (reference order) of a leaf, -1 for a twig.
 * Vectors are [cap][3] for xx, lx, D, Q, QQ.  *nnodes receives the count
 * (TB_ERR_ARG if cap is too small).
 */
int tb_export_tree(tb_handle *h, int32_t species, int32_t cap, int32_t *nnodes, int32_t *depth, int32_t *leaf,
                   int32_t *pt, int32_t *count, int32_t *next, double *xx, double *lx, double *qeff, double *dx,
                   double *D, double *Q, double *QQ);

/*
 * FP64 O(N^2) direct sum on sampled targets with the reference's pair rules
 * (like species Coulomb with r^2 > 0.01 um^2, unlike species simplified
 * Kelbg).  tspecies[t] 0/1, tidx[t] index in reference order; f out [nt][3].
 */
int tb_direct_sample(tb_handle *h, int32_t nt, const int32_t *tspecies, const int32_t *tidx, double *f);

/*
 * tb_direct_sample without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * target list and the forces are staged in a fixed per-GPU array of the
 * library (2.25 MiB, held until the process ends) instead of memory allocated
 * per call, so nt is at most TB_NOALLOC_MAX_TARGETS (TB_ERR_ARG beyond).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (handles on the same device driven from several host threads) race.  Same
 * forces as tb_direct_sample, bit for bit.
 */
#define TB_NOALLOC_MAX_TARGETS 65536
int tb_direct_sample_noalloc(tb_handle *h, int32_t nt, const int32_t *tspecies, const int32_t *tidx, double *f);

/*
 * Check of the mixed walk's far-field local expansion (no reference
 * counterpart; tests only).  cells [ncell][28]: R = expansion centre - cell
 * centre (3), q, D (3), Q'xx, Q'yy, Q'xy, Q'yz, Q'zx (traceless quadrupole),
 * O'xxx, O'xxy, O'xxz, O'xyy, O'xyz, O'yyy, O'yyz (traceless third moment;
 * orders 4 to 6), H'xxxx, H'xxxy, H'xxxz, H'xxyy, H'xxyz, H'xyyy, H'xyyz,
 * H'yyyy, H'yyyz (traceless fourth moment; order 6, the walk's).  The cells'
 * local coefficients of order `order` (3 to 6)
 * are summed across a wavefront as in the walk and evaluated at the offsets
 * y [ny][3]: out [ny][4] = potential, field x, y, z (FP32).
 */
int tb_local_eval(tb_handle *h, int32_t order, int32_t ncell, const float *cells, int32_t ny, const float *y,
                  float *out);

/*
 * The FP64 walk (TB_WALK_REF64 trees) for a list of targets only (reference
 * order): f [nt][3].  Same forces as tb_force gives those targets.
 */
int tb_force_targets(tb_handle *h, int32_t nt, const int32_t *tspecies, const int32_t *tidx, double *f);
/* device time of the FP64 walk over every stride-th block of 64 consecutive
 * (key-ordered) targets of each species, so that each wavefront walks
 * neighbouring targets as in the full walk (sampled timing; per-target walks
 * are independent); *nwalked receives the number of targets walked */
int tb_time_walk(tb_handle *h, int32_t stride, float *ms, int64_t *nwalked);

typedef struct tb_stats {
    int64_t nodes[2];        /* node count per species tree                     */
    int32_t max_depth[2];
    int32_t pad_;
    int64_t visits[4];       /* MAC tests   ee, ie, ei, ii (last counted force) */
    int64_t m2p[4];          /* accepted cells                                  */
    int64_t p2p[4];          /* leaf interactions                               */
    double ms_build, ms_force, ms_vv;  /* last call, device time               */
    double ms_walk;          /* last force: the walk kernel alone (mixed walk;  */
                             /* the FP64 walk reports its force phase)          */
    int64_t m2l[4];          /* mixed walk: far cells taken through the target  */
                             /* group's local expansion (counted per target)    */
    int64_t far_pairs[2];    /* the same as (far cell, target group) pairs, by  */
                             /* target species                                  */
    int64_t seg_pairs[2];    /* two-level far field: (far cell, segment) pairs,  */
                             /* by target species (their (cell, target) pairs   */
                             /* are in m2l)                                     */
    int64_t m2p3[2];         /* mixed walk: of the m2p, the (cell, target) pairs */
                             /* evaluated to the third moment only (cells far   */
                             /* from the group for their size), by target       */
                             /* species                                         */
} tb_stats;
/* counting pass: re-runs the last force with interaction counters */
int tb_count_interactions(tb_handle *h, tb_stats *s);
int tb_get_stats(tb_handle *h, tb_stats *s);

/* ------------------------------------------------ learned opening criterion
 * With mac = TB_MAC_LEARNED the mixed walk accepts every cell beyond
 * s/theta_lo + dx, opens every cell within s/theta_mixed + dx, and in between
 * asks the MLP (8 -> 32 -> 32 -> 1, BF16 weights, FP32 accumulation) for the
 * predicted log10 relative force error of accepting the cell for the whole
 * target group; it accepts when that is below log10(err_target).  Without a
 * trained network it falls back to the fixed-theta walk (theta_mixed).
 * No counterpart in the reference (its criterion is the fixed theta of
 * tree.f90:742).
 */
#define TB_MLP_NPARAMS 1377
int tb_mlp_set(tb_handle *h, const float *params, int32_t n);
int tb_mlp_get(tb_handle *h, float *params, int32_t n);
/* seeded initial network (He-uniform weights, zero biases) for tb_mlp_train */
int tb_mlp_init(tb_handle *h, uint64_t seed);
/* forward pass on n samples X[n][8] (bf16 != 0: weights rounded to BF16) */
int tb_mlp_forward(tb_handle *h, int32_t n, const float *X, int32_t bf16, float *out);
/* mean squared error and its gradient (fp32 master weights) */
int tb_mlp_loss_grad(tb_handle *h, int32_t n, const float *X, const float *y, float *loss, float *grad);
/* Adam on the GPU; loss_hist (optional) gets the per-step batch loss */
int tb_mlp_train(tb_handle *h, int32_t n, const float *X, const float *y, int32_t steps, int32_t batch, float lr,
                 uint64_t seed, float *loss_hist);
/*
 * Training data from the method itself: runs the mixed walk accepting cells
 * beyond s/theta_lo + dx, records each (cell, group) pair whose acceptance
 * the learned criterion would decide, with its 8 features X[k][8], and labels
 * it y[k] = log10 rms over the group of |F_M2P - F_exact| / |F| (FP64).
 * One fixed-theta force evaluation runs first: one feature is the group's
 * smallest net field from the previous force.
 */
int tb_mlp_collect(tb_handle *h, int32_t cap, float *X, float *y, int32_t *n);

/* time the mixed walk on the current tree for 64, 128 and 256 targets per
 * group, each with interaction lists evaluated at 15 / 11 / 7 cells (24 / 16 /
 * 12 leaves), keep the fastest pair for this problem size; ms_out[3]
 * (optional): the three group sizes at the chosen list lengths */
int tb_autotune(tb_handle *h, int32_t *group_out, float *ms_out);
/* the list lengths (cells, leaves) the mixed walk of this handle flushes at */
int tb_walk_lists(tb_handle *h, int32_t *flush_m, int32_t *flush_p);

/* ------------------------------------------------------- batched ensemble
 * Many independent realisations (each its own electron and ion trees) on one
 * GPU: per step the builds of all realisations share one host read-back and
 * the force is one persistent walk launch over all their target groups, with
 * the autotuned group size.  Mixed walk only.  No counterpart in the
 * reference, which runs one realisation per program run.
 */
typedef struct tb_ensemble tb_ensemble;
int tb_ensemble_create(tb_ensemble **e, int32_t n, const tb_config *cfgs);
int tb_ensemble_destroy(tb_ensemble *e);
int tb_ensemble_set_particles(tb_ensemble *e, int32_t r, int32_t ne, const tb_pt *el, int32_t ni, const tb_pt *io);
int tb_ensemble_get_particles(tb_ensemble *e, int32_t r, tb_pt *el, tb_pt *io);
/* build + force for all; epot (optional) [n][2] = Epot_e, Epot_i */
int tb_ensemble_force(tb_ensemble *e, double *epot);
/* nsteps for all; out (optional) [n][6] as tb_step after the last step */
int tb_ensemble_step(tb_ensemble *e, double dt, int32_t nsteps, double *out);
int tb_ensemble_autotune(tb_ensemble *e, int32_t *group_out, float *ms_out);
int tb_ensemble_walk_lists(tb_ensemble *e, int32_t *flush_m, int32_t *flush_p);

/* ------------------------------------------------ radial distribution
 * Leaf-pair RDF histograms, as the reference accumulates them in every
 * LocalCoulombSum (tree.f90:916-919, 1021-1024, 1125-1129; NS%RDF_ei /
 * NS%RDF_ii, data.f90:139) and RDF prints and resets them (sample.f90:5-50):
 * every unlike leaf pair of the ie and ei walks adds 1 to rdf_ei[INT(r/dx_ei)],
 * every like ion pair beyond the 0.1 um cut adds 1 to rdf_ii[INT(r/dx_ii)],
 * for bins 0 .. nbin-1 (the reference's 1 .. Ndist+1, Ndist = 10000).  The
 * FP64 walk accumulates them (TB_WALK_REF64 only); the counting pass adds
 * nothing.  nbin = 0 disables and frees them.
 */
int tb_rdf_enable(tb_handle *h, double dx_ei, double dx_ii, int32_t nbin);
/* copy out (and with reset != 0 clear) the histograms accumulated since enable / the last reset */
int tb_rdf_get(tb_handle *h, int32_t nbin, int64_t *rdf_ei, int64_t *rdf_ii, int32_t reset);

const char *tb_last_error(void);

/* sizeof(tb_pt), sizeof(tb_config), sizeof(tb_stats) as compiled (binding checks) */
int tb_abi_sizes(int32_t *pt, int32_t *config, int32_t *stats);

/* ------------------------------------------------------------ host helpers */

/* Kelbg length: thermal de Broglie lambda hbar/sqrt(2 mu k_B T_e), um */
double tb_kelbg_length(double te_kelvin, double xm_e, double xm_i);

/*
 * Seeded synthetic plasma: spherical Gaussian positions (sigma um) per species
 * from Philox4x32-10 + Box-Muller, Maxwell-Boltzmann velocities rescaled to
 * exactly T (Vinit, main.f90:709-767: 3N-6 degrees of freedom).  ids 1..n.
 */
int tb_gen_plasma(uint64_t seed, int32_t ne, int32_t ni, double sigma, double te, double ti, double xm_e,
                  double xm_i, tb_pt *e, tb_pt *i);

#ifdef __cplusplus
}
#endif
#endif
