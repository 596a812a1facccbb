/*
 * yaff_nb.h -- C ABI of the MI355X nonbonded engine (libyaff_nb.so).
 *
 * Every entry point is asynchronous on the HIP stream it is given, takes
 * device pointers (allocated by the caller) plus sizes, allocates nothing and
 * never synchronises, except those that say so: ynb_nlist_rows
 * (synchronises) and the host-only planning functions (ynb_sys_plan,
 * ynb_ewald_kvectors, ynb_row_blocks, ynb_ewald_records, ynb_ewald_work).
 * Return value: 0 on success, a negative YNB_E* code otherwise;
 * ynb_last_error() gives the message.
 *
 * Work is batched: a launch covers every system of a ynb_batch.  Atoms of
 * system s live at [sys.atom0, sys.atom0 + sys.natom) of the padded atom
 * arrays; atom0 is a multiple of 256 so that neither a wavefront nor the
 * atom blocks of the Ewald force pass (up to 256 atoms) hold atoms of two
 * systems.  A single system is a batch of one.  A batch may hold free
 * ranges between its systems (wavefronts, bins, k-vector blocks and
 * structure-factor work items whose owner entry is -1, skipped by every
 * kernel) and free systems (natom 0, no potentials, nk 0): a persistent
 * batch admits and retires systems by rewriting these tables in place.
 *
 * Reference interfaces replaced (molmod/yaff):
 *   ynb_nlist_build + ynb_nlist_rows   <- nlist_build_low   yaff/pes/nlist.h:37-39
 *                                         (+ nlist_status_* yaff/pes/ext.pyx:359-449)
 *   ynb_nlist_recompute                <- nlist_recompute_low yaff/pes/nlist.h:41-42
 *   ynb_pair_compute, ynb_pair_rows    <- pair_pot_compute  yaff/pes/pair_pot.h:61-64
 *                                         with pair_fn_lj/mm3/exprep/ei and
 *                                         switch3/hammer (yaff/pes/truncation.h)
 *   ynb_ewald_reci                     <- compute_ewald_reci yaff/pes/ewald.h:31-34
 *   ynb_ewald_corr                     <- compute_ewald_corr yaff/pes/ewald.h:39-43
 *                                         (batched: every system of a ynb_batch)
 *   ynb_verlet_*                       <- VerletIntegrator.propagate
 *                                         yaff/sampling/verlet.py:339-367
 *   ynb_ewald_kvectors                 <- the k loop of compute_ewald_reci
 *                                         yaff/pes/ewald.c:38-58 (host planning)
 * No counterpart in the reference (batches and their tables):
 *   ynb_row_blocks, ynb_ewald_records, ynb_ewald_work(_device), ynb_scatter,
 *   ynb_total, ynb_pair_rows_static
 */
#ifndef YAFF_NB_H
#define YAFF_NB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *ynb_stream; /* hipStream_t */

#define YNB_MAX_POT 4
#define YNB_WAVE 64

enum { YNB_OK = 0, YNB_EINVAL = -1, YNB_EHIP = -2, YNB_EWORKSPACE = -3, YNB_ECAPACITY = -4 };
enum { YNB_POT_LJ = 0, YNB_POT_MM3 = 1, YNB_POT_EXPREP = 2, YNB_POT_EI = 3 };
enum { YNB_TR_NONE = 0, YNB_TR_SWITCH3 = 1, YNB_TR_HAMMER = 2 };
enum { YNB_FP64 = 0, YNB_MIXED = 1 };

/* Same layout as neigh_row_type (yaff/pes/nlist.h:30-35): 72 bytes. */
typedef struct ynb_neigh_row {
  int64_t a, b;
  double d, dx, dy, dz;
  int64_t r0, r1, r2;
} ynb_neigh_row;

/* Same layout as scaling_row_type (yaff/pes/pair_pot.h:46-50). */
typedef struct ynb_scaling_row {
  int64_t a, b;
  double scale;
  int64_t nbond;
} ynb_scaling_row;

/* One pair potential of a system (one ForcePartPair). */
typedef struct ynb_pot {
  int32_t kind;       /* YNB_POT_* */
  int32_t trunc;      /* YNB_TR_* */
  double rcut;
  double trunc_par;   /* Switch3 width or Hammer tau */
  double alpha;       /* EI: Ewald alpha (0: bare Coulomb) */
  double dielectric;  /* EI */
  int32_t nffatype;   /* ExpRep: number of atom types */
  int32_t cross0;     /* ExpRep: offset of the amp table in ynb_pairpar.cross; b follows */
  int32_t stab0;      /* scaling rows of this potential: [stab0, stab0+nstab) of ynb_pairpar.stab */
  int32_t nstab;
} ynb_pot;

/* One system of a batch.  Filled on the host (ynb_sys_plan completes the
 * neighbour-list grid) and copied to the device by the caller. */
typedef struct ynb_sys {
  double rvecs[9];    /* cell vectors as rows (full 3x3, Cell.update_rvecs) */
  double gvecs[9];    /* reciprocal rows, no 2 pi */
  double rc;          /* neighbour-list radius: rcut + skin */
  double volume;
  int32_t nvec;       /* 0..3 periodic directions */
  int32_t natom;
  int32_t atom0;      /* offset into the padded atom arrays, multiple of 256 */
  int32_t npot;
  int32_t rmax[3];    /* image range of nlist.py:132 */
  int32_t nlow, nhigh;
  int32_t grid[3];    /* bins per cell vector (ynb_sys_plan) */
  int32_t range[3];   /* bins searched on each side (ynb_sys_plan) */
  int32_t bin0;       /* first global bin of this system */
  int32_t nbin;
  int32_t any_stab;   /* nonzero when a potential of the system has scaling rows */
  int32_t box_safe;   /* ynb_sys_plan: Yaff's image box cannot reject a pair within rc */
  ynb_pot pot[YNB_MAX_POT];
  /* Ewald reciprocal part (nk == 0: none) */
  double alpha, dielectric;
  int32_t k0, nk;         /* k vectors [k0, k0+nk) of ynb_ewald.kvec; k0 and nk multiples of 256 */
  int32_t frb0, nfrb;     /* row blocks of the force pass: [frb0, frb0+nfrb) of ynb_ewald.fblk */
  int32_t nslot;          /* partial-gradient slots of the force pass: ceil(nfrb / 4) */
  int32_t ew_nlow;        /* atom groups of the Ewald reciprocal part (ForcePartEwaldReciprocal nlow/nhigh,
                             yaff/pes/ff.py:331-401): its own, independent of the neighbour list's nlow/nhigh */
  int32_t chunk0, nchunk; /* atom chunks of the structure-factor pass (64 * apl atoms each) */
  int64_t spart0;         /* offset of this system's partial structure factors, in k entries */
  int32_t nchunk_low, nchunk_high; /* chunks [0, low) below ew_nlow, the last `high` at or above ew_nhigh */
  int32_t pair_form;      /* YNB_FORM_*: specialised pair kernel for this system's potential set */
  int32_t ew_nhigh;
  double ew_const;        /* constant added to the energy of the system's Ewald reciprocal out: the self term
                             of its Ewald correction (ewald.c:286-290, times 1/dielectric) when the system has
                             no scaled pairs, so that no ynb_ewald_corr launch is needed for it; else 0 */
} ynb_sys;

/* ynb_sys.pair_form: potential sets with a specialised totals kernel (no
 * scaling rows; ynb_pairpar.rec holds the per-atom record).  Any other set,
 * per-potential results (gpos_slots) or a slot mask that leaves out one of
 * the system's potentials use the generic kernel. */
#define YNB_FORM_GENERIC 0
#define YNB_FORM_LJ 1     /* one LJ, no truncation or Switch3; rec = (sigma, sqrt(eps), 0, 0) */
#define YNB_FORM_IONIC 2  /* ExpRep, LJ, EI (alpha > 0, point charges), no truncation, in this order;
                             rec = (LJ sigma, sqrt(LJ eps), charge / sqrt(dielectric), ExpRep type) */

/* Device tables shared by all launches over one batch. */
typedef struct ynb_batch {
  const ynb_sys *sys;        /* device [nsys] */
  const int32_t *wave_sys;   /* device [npad/64]: system of each wavefront, -1 = none */
  const int32_t *bin_sys;    /* device [nbin]: system of each bin, -1 = none (a free range) */
  int32_t nsys;
  int32_t nbin;              /* total bins (0 when every system uses the all-pairs path) */
  int64_t npad;              /* padded atom count, multiple of 256 */
  int32_t max_bin_atoms;     /* capacity of a bin, 1..1024 (ynb_nlist_build flags overflow) */
  int32_t flags;             /* YNB_BATCH_*: which build paths the batch needs */
} ynb_batch;

/* ynb_batch.flags */
#define YNB_BATCH_ALLPAIRS 1    /* some system has natom > 0 and no bins (nvec < 3 or a huge search box) */
#define YNB_BATCH_EXACT_BUILD 2 /* some binned system is not box_safe: exact re-test inside the build */
#define YNB_BATCH_FORM_GENERIC 4 /* some system has pair_form YNB_FORM_GENERIC */
#define YNB_BATCH_FORM_LJ 8      /* some system has pair_form YNB_FORM_LJ */
#define YNB_BATCH_FORM_IONIC 16  /* some system has pair_form YNB_FORM_IONIC */
#define YNB_BATCH_WIDE_SEARCH 32 /* some binned system stages more than YNB_BUILD_NB_SMALL partner
                                    bins per home group (ynb_build_search_bins) */
#define YNB_BUILD_NB_SMALL 160   /* partner bins of the build's small tables (two 5x5x5 search boxes
                                    one bin apart: 150) */

/* Full neighbour list, one row of kcap entries per atom: entry k of atom i is
 * at i*kcap + k.
 * j is the padded index of the partner, img packs the integer translation n
 * (10 bits per axis, offset 512) with pos_j + n.R the partner image. */
typedef struct ynb_nlist {
  int32_t *j;
  int32_t *img;
  int32_t *cnt;        /* [npad] entries per atom (also counted past kcap) */
  int32_t *status;     /* [4]: [0] flags: 1 row > kcap, 2 image shift > 511 cells, 4 bin > max_bin_atoms,
                          8 partner bins > the tables (YNB_BATCH_WIDE_SEARCH missing);
                          [1] after flag 4, the fullest bin's atom count.
                          Flags are sticky (builds OR into them, the caller clears them), so
                          a check after many graph-replayed steps sees every overflow.  On a
                          flag the caller reads cnt (rows) or status[1] (bins) to size the
                          retry. */
  const int32_t *rebuild; /* [nsys] per-system rebuild flags (ynb_nlist_skin_check) or NULL: rebuild all */
  int32_t kcap;
  int32_t flags;       /* YNB_NLIST_*: 0, or PREBINNED when a ynb_verlet_kick_drift given this list and
                          workspace has already binned the atoms at the positions of this build */
} ynb_nlist;

#define YNB_NLIST_PREBINNED 1  /* (requires rebuild == NULL) */

/* Per-atom parameters of the pair potentials. */
typedef struct ynb_pairpar {
  const double *par;   /* [YNB_MAX_POT][3][npad]: LJ/MM3 sigma, epsilon, onlypauli;
                          EI charge, radius; ExpRep ffatype id (as double) */
  const double *cross; /* ExpRep amp/b tables, row-major nffatype^2 each */
  const ynb_scaling_row *stab; /* scaling rows (a > b, local indices), sorted per system */
  const double *rec;   /* [npad][4] per-atom record of the specialised forms (YNB_FORM_*) */
  int32_t *ctr;        /* [nsys] block counters of ynb_pair_compute (zero-initialised; left at zero) */
  int32_t lanes;       /* lanes per atom in ynb_pair_compute: 16, 32 or 64 (0: 16) */
  int32_t pad_;
} ynb_pairpar;

/* Ewald reciprocal-space tables and work buffers (all device).
 *
 * The k vectors of a system form rows (runs of consecutive g0 at fixed g1,
 * g2, in the order g2, g1 major, g0 ascending).  Both passes walk blocks of
 * whole rows: a lane holds atoms, seeds exp(i k.r) at each row start and
 * steps it along g0.  The structure-factor pass covers every (row block, atom
 * chunk) work item; the force pass gives each wavefront one row block of a
 * system and 64 * fapl atoms, and a block of 4 wavefronts sums its 4 row
 * blocks. */
typedef struct ynb_ewald {
  const double *kvec;     /* [nk_tot][4]: kx, ky, kz and c_k = 4 pi/V exp(-k^2/4 alpha^2)/k^2 (0: padding) */
  const int32_t *krow;    /* [nrow_tot][4]: g0 of the row's first k, g1, g2, number of k in the row */
  const int32_t *krow_k;  /* [nrow_tot]: global index of the row's first k */
  const int32_t *skblk;   /* [nskblk][2]: global row range [begin, end) of each structure-factor block */
  const int32_t *fblk;    /* [nfblk][2]: global row range of each force block (system s: ynb_sys.frb0..) */
  const int32_t *chunk;   /* [nchunk_tot][4]: system, first local atom, atom count, group 0 mid/1 low/2 high */
  const int32_t *work;    /* [nwork][4]: system (-1: skip), structure-factor block, global chunk, 0 */
  const int32_t *kblk_sys;/* [nk_tot/256]: system of each block of 256 k vectors, -1 = none */
  const int32_t *kpos;    /* [nk_tot]: force record of each k (-1: padding k) */
  const int32_t *krow_f;  /* [nrow_tot]: first force record of each row */
  int32_t *fctr;          /* [npad/64] block counters of the force pass (zero-initialised; left at zero) */
  int32_t nwork;
  int32_t nk_tot;
  int32_t apl;            /* atoms per lane of the structure-factor pass: 2, 4 or 8 */
  int32_t fslots;         /* partial-gradient slots: the largest ynb_sys.nslot */
  int32_t groups;         /* nonzero when some system has nlow > 0 or nhigh < natom */
  int32_t fapl;           /* atoms per lane of the force pass: 1, 2 or 4 */
  int32_t flags;          /* YNB_EWALD_*: 0, or PREPARED when a ynb_verlet_kick_drift given this
                             struct has already filled frac and step for the current positions */
  int32_t pad_;
  double *frac;           /* work [4][npad]: wrapped fractional coordinates and charge */
  double *step;           /* work [4][npad]: cos, sin of 2 pi f0, then of 2 pi f1 */
  double *spart;          /* work: partial structure factors, 2 doubles per (chunk, k) */
  double *coef;           /* work [nk_tot][6]: 2 c_k S(k) of the mid, low and high groups (cos, sin) */
  double *frec;           /* work [nfrec][2]: 2 c_k S_sin, 2 c_k S_cos of the mid group per k; each row
                             padded to a multiple of 4 records, padding and 4 spare records at the end zero */
  double *kpart;          /* work [nk_tot/32][8] */
  double *gpart;          /* work [fslots][npad][3] */
} ynb_ewald;

#define YNB_EWALD_PREPARED 1

/* ------------------------------------------------------------------ */
const char *ynb_version(void);
const char *ynb_last_error(void);

/* Kernel timing: while enabled, each kernel launched by the library is
 * bracketed by two HIP events (created once, on the first enable) recorded on
 * its own stream.  ynb_profile_read synchronises on launch i's end event and
 * returns its kernel name and duration.  Enabling resets the record. */
int ynb_profile_enable(int on);
int ynb_profile_count(void);
int ynb_profile_read(int i, const char **name, float *ms);

/* Host only.  Completes sys->grid, sys->range and sys->nbin from the cell,
 * sys->rc and sys->natom.  Systems with nvec < 3 use the all-pairs path
 * (nbin = 0). */
int ynb_sys_plan(ynb_sys *sys);

/* Host only.  Partner bins the cell-binned build stages per home group of a
 * planned system (a group: up to two bins adjacent along the third cell
 * vector whose search boxes are staged together); a batch holding a system
 * above YNB_BUILD_NB_SMALL needs YNB_BATCH_WIDE_SEARCH.  0 for a system
 * without bins. */
int ynb_build_search_bins(const ynb_sys *sys);

/* Host only.  The k vectors of compute_ewald_reci (yaff/pes/ewald.c:38-58)
 * for reciprocal rows gvecs [3][3] (no 2 pi), cell volume, alpha, gcut and
 * the image range gmax[3], with the same FP64 expressions for k and |k|^2
 * (hence the same set), in row order (g2, then g1, then g0 ascending):
 * kvec [nk][4] = kx, ky, kz, c_k = 4 pi/V exp(-k^2/4 alpha^2)/k^2; g [nk][3];
 * rows, the runs of consecutive g0 at fixed (g1, g2): krow [nrow][4] = first
 * g0, g1, g2, count and krow_k [nrow] = the row's first k.  Every output
 * holds cap entries (nrow <= nk); YNB_ECAPACITY when more k vectors pass. */
int ynb_ewald_kvectors(const double *gvecs, double volume, double alpha, double gcut, const int64_t *gmax,
                       int64_t cap, double *kvec, int32_t *g, int32_t *krow, int32_t *krow_k, int64_t *nk,
                       int64_t *nrow);

/* Host only.  Consecutive rows (krow as above) grouped greedily into blocks
 * of at most kmax k vectors (a longer row is a block of its own): blk
 * [nblk][2] = first row, end row; blk holds nrow entries. */
int ynb_row_blocks(const int32_t *krow, int64_t nrow, int32_t kmax, int32_t *blk, int64_t *nblk);

/* Host only.  Force records of one system's rows (krow as above): each
 * row padded to a multiple of 4 records; krow_f [nrow] = the row's first
 * record, kpos [nkp] = the record of each k (-1 for k >= nk), *nrec the
 * total (ynb_ewald.krow_f, kpos and frec before the batch offsets). */
int ynb_ewald_records(const int32_t *krow, int64_t nrow, int64_t nkp, int32_t *krow_f, int32_t *kpos,
                      int64_t *nrec);

/* Host only.  The structure-factor work items of one system, every (row
 * block, atom chunk), block major: work [nblk * nchunk][4] = sys, blk0 +
 * block, chunk0 + chunk, 0 (ynb_ewald.work). */
int ynb_ewald_work(int64_t nblk, int64_t nchunk, int32_t sys, int32_t blk0, int32_t chunk0, int32_t *work);
/* The same items written by a kernel into device memory (work: the
 * [nblk * nchunk][4] slice of ynb_ewald.work, 16-byte aligned), on stream. */
int ynb_ewald_work_device(int32_t *work, int64_t nblk, int64_t nchunk, int32_t sys, int32_t blk0, int32_t chunk0,
                          ynb_stream stream);

/* Workspace of ynb_nlist_build in bytes, for bins of bin_cap atoms
 * (ynb_batch.max_bin_atoms).  The caller zero-initialises it once; the build
 * leaves its bin counters at zero.  The counters are its first 4 * nbin
 * bytes, so a batch whose npad or nbin changes keeps them; when nbin grows,
 * the caller zeroes the counters of the new bins. */
size_t ynb_nlist_workspace_bytes(int64_t npad, int32_t nbin, int32_t bin_cap);

/* Workspace of ynb_nlist_rows for a system of natom atoms and nrows rows. */
size_t ynb_rows_workspace_bytes(int64_t natom, int64_t nrows);

/* Cell-binned (or all-pairs) neighbour list of every system: the full list of
 * partners j of each atom i (both directions, periodic self images included)
 * whose Yaff half row -- built exactly as nlist_build_low builds it -- has
 * d^2 < rc^2.  Overflows are flagged in nl->status (see ynb_nlist): the
 * caller grows kcap or max_bin_atoms and builds again.  Uses no memset, no
 * library sort and no synchronisation, so it can be captured in a graph. */
int ynb_nlist_build(const ynb_batch *b, const double *pos, void *work, size_t work_bytes,
                    const ynb_nlist *nl, ynb_stream stream);

/* Verlet skin: rebuild[s] = 1 when skin[s] <= 0 or some atom of system s
 * moved by skin[s]/2 or more since pos_ref (then pos_ref of that system is
 * set to pos); 0 otherwise.  Every pair within rcut = rc - skin then stays in
 * the list built with rc: the entries carry fixed image translations, so a
 * pair distance changes by at most the sum of the two displacements.
 * (Yaff's row lists re-image with the minimum image, hence the factor
 * 2 (rmax + 1) of yaff/pes/nlist.py:207-218.)  pos_ref starts as a huge
 * value so that the first test rebuilds. */
int ynb_nlist_skin_check(const ynb_batch *b, const double *pos, double *pos_ref, const double *skin,
                         int32_t *rebuild, ynb_stream stream);

/* Materialises system `s`'s half list as Yaff rows (local atom indices) in
 * nlist_build_low's order.  Synchronises.  Writes at most `cap` rows to
 * `rows` (device) and the total to *count (host). */
int ynb_nlist_rows(const ynb_batch *b, int32_t s, const double *pos, const ynb_nlist *nl,
                   void *work, size_t work_bytes, ynb_neigh_row *rows, int64_t cap,
                   int64_t *count, ynb_stream stream);

/* Recomputes d, dx, dy, dz of existing rows from new positions (skin path). */
int ynb_nlist_recompute(const ynb_sys *h_sys, const double *pos, const double *pos_old,
                        ynb_neigh_row *rows, int64_t nrow, ynb_stream stream);

/* Pair potentials of the slots in slot_mask, one pass over the list.
 * gpos (nullable): [npad][3] gradient summed over the slots, written or, with
 * accumulate != 0, added.  gpos_slots (nullable): [YNB_MAX_POT][npad][3].
 * out: [nsys][YNB_MAX_POT][10] energy followed by the 3x3 virial.
 * partial: work of npad*lanes/256 * YNB_MAX_POT * 8 doubles; pp->ctr: the
 * per-system block counters (the last block of a system reduces its sums). */
int ynb_pair_compute(const ynb_batch *b, const ynb_pairpar *pp, const double *pos,
                     const ynb_nlist *nl, uint32_t slot_mask, int32_t precision,
                     double *gpos, int32_t accumulate, double *gpos_slots,
                     double *partial, double *out, ynb_stream stream);

/* pair_pot_compute over explicit Yaff rows (local atom indices) for one
 * potential (yaff/pes/pair_pot.h:61-64).  Parameter arrays are per atom
 * (p1/p2/cross as in ynb_pairpar, nullable when unused).  out[10] = energy
 * and, with derivs != 0, the virial (zeros otherwise).  With derivs and gpos
 * (nullable, [natom][3]) the gradient is ADDED to gpos: each atom sums its
 * rows in the order atom_ref lists them -- atom_ref0 [natom + 1] offsets,
 * atom_ref the row references of each atom, 2 r where it is row r's b (+)
 * and 2 r + 1 where it is a (-), in ascending row order (b before a for a
 * self-image row), which is the order of pair_pot.c:147-157 -- so the result
 * is deterministic and needs no atomics.  Work: rowh [nrow][3] (with
 * gradients) and partial [max(1, ceil(nrow / 256))][8].  nrow <= 2^30. */
int ynb_pair_rows(const ynb_neigh_row *rows, int64_t nrow, const ynb_scaling_row *stab, int64_t nstab,
                  const ynb_pot *pot, const double *p0, const double *p1, const double *p2,
                  const double *cross, int32_t derivs, int64_t natom, const int32_t *atom_ref0,
                  const int32_t *atom_ref, double *gpos, double *rowh, double *partial, double *out,
                  ynb_stream stream);

/* ynb_pair_rows with no work arguments: rowh and partial live in a static
 * device array of the library (one per GPU, held until the process ends),
 * so the call allocates nothing, for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  At most YNB_PAIR_ROWS_STATIC_MAX_ROWS rows
 * (YNB_ECAPACITY beyond: about 6 MB of device memory per GPU).  All calls on
 * one GPU share the array: calls that overlap on one GPU (other streams,
 * other host threads) race and must be serialised by the caller. */
#define YNB_PAIR_ROWS_STATIC_MAX_ROWS 262144
int ynb_pair_rows_static(const ynb_neigh_row *rows, int64_t nrow, const ynb_scaling_row *stab, int64_t nstab,
                         const ynb_pot *pot, const double *p0, const double *p1, const double *p2,
                         const double *cross, int32_t derivs, int64_t natom, const int32_t *atom_ref0,
                         const int32_t *atom_ref, double *gpos, double *out, ynb_stream stream);

/* Ewald reciprocal term of every system with nk > 0.
 * out: [nsys][10] energy followed by the virial (zeros for systems without k
 * vectors); gpos as in ynb_pair_compute.  The partial gradients of each
 * system's force slots are left in ew->gpart ([slot][npad][3], ynb_sys.nslot
 * slots); with gpos non-NULL the last force block of each atom block adds
 * them into gpos (counters ew->fctr). */
int ynb_ewald_reci(const ynb_batch *b, const ynb_ewald *ew, const double *pos, const double *charges,
                   int32_t precision, double *gpos, int32_t accumulate, double *out,
                   ynb_stream stream);

/* Ewald correction (ForcePartEwaldCorrection) of every system of a batch:
 * the self term and the erf term of the scaled pairs.  Per system: */
typedef struct ynb_corr_sys {
  double alpha, dielectric;
  double self_energy;     /* -(alpha / sqrt(pi)) sum over nlow <= i < nhigh of q_i^2 (constant; the host
                             sums it in atom order, as ewald.c:286-290) */
  int32_t nlow, nhigh;    /* the part's atom groups (local indices) */
  int32_t row0, nrow;     /* its scaling rows: [row0, row0 + nrow) of ynb_corr.stab */
  int32_t blk0, nblk;     /* its row blocks of 256 rows: [blk0, blk0 + nblk), nblk >= 1 */
} ynb_corr_sys;

/* Tables of a batch (device).  atom_ref lists, for each padded atom in
 * ascending row order, the rows that hold it: 2 r when it is row r's b (the
 * gradient term is added), 2 r + 1 when it is row r's a (subtracted); so
 * each atom's gradient is summed in the order in which compute_ewald_corr
 * accumulates it, without atomics. */
typedef struct ynb_corr {
  const ynb_corr_sys *csys;     /* [nsys] */
  const ynb_scaling_row *stab;  /* [nrow_tot] (a > b, local indices), each system's rows sorted as Scalings.stab */
  const int32_t *blk_sys;       /* [nblk_tot]: system of each row block */
  const int32_t *atom_ref0;     /* [npad + 1]: offsets of each atom's references in atom_ref */
  const int32_t *atom_ref;      /* [2 nrow_tot] */
  double *rowg;                 /* work [nrow_tot][3]: delta * g of each row (before 1/dielectric) */
  double *partial;              /* work [nblk_tot][8]: energy and virial sums of each row block */
  int32_t *ctr;                 /* [nsys] block counters (zero-initialised; left at zero) */
  int32_t nrow_tot, nblk_tot;
} ynb_corr;

/* ynb_ewald_corr replaces compute_ewald_corr (yaff/pes/ewald.h:39-43) for
 * every system of the batch at once.  out: [nsys][10] energy followed by the
 * virial, written or, with out_accumulate, added (the Verlet step adds the
 * correction to the reciprocal term's out this way); gpos (nullable)
 * [npad][3] written or, with accumulate, added, for every atom of a system
 * (atoms without scaled pairs get zero).  Two launches: the rows (energy,
 * virial, delta * g per row; each system's last row block sums its block
 * sums in a fixed order), then the atoms (each sums its rows in row order).
 * Deterministic. */
int ynb_ewald_corr(const ynb_batch *b, const ynb_corr *c, const double *pos, const double *charges,
                   double *gpos, int32_t accumulate, double *out, int32_t out_accumulate, ynb_stream stream);

/* Velocity Verlet halves (yaff/sampling/verlet.py:345-357): kick_drift does
 * vel += dt/2 * (-gpos/m); pos += dt * vel; kick does vel += dt/2 * (-gpos/m)
 * and writes 2*Ekin of each system to ekin2[nsys] via `partial`
 * (npad/64 doubles; the last block of each system adds them, counters
 * ctr[nsys] zero-initialised and left at zero).  gpart (NULL or the ynb_ewald
 * gpart of a ynb_ewald_reci call made with gpos = NULL): kick first adds each
 * system's nslot Ewald partial gradients to gpos, in slot order, and stores
 * the sum in gpos -- the Ewald gradient sum folded into the kick.  total
 * (NULL or as in ynb_total, with pair_out/ewald_out nullable): kick also
 * totals each system's energy and virial, in place of a ynb_total call.
 * kick_drift with ew (nullable; charges as in ynb_ewald_reci) also fills
 * ew->frac and ew->step for the new positions, after which ynb_ewald_reci
 * with ew->flags = YNB_EWALD_PREPARED skips that pass; with nl_work (the
 * ynb_nlist_build workspace, nullable) and nl (rebuild == NULL) it also bins
 * the atoms at the new positions, after which ynb_nlist_build with
 * nl->flags = YNB_NLIST_PREBINNED skips its binning pass. */
int ynb_verlet_kick_drift(const ynb_batch *b, double *pos, double *vel, const double *gpos,
                          const double *minv, const double *dt, const ynb_ewald *ew,
                          const double *charges, void *nl_work, size_t nl_work_bytes,
                          const ynb_nlist *nl, ynb_stream stream);
int ynb_verlet_kick(const ynb_batch *b, double *vel, double *gpos, const double *gpart,
                    const double *minv, const double *masses, const double *dt, double *partial,
                    int32_t *ctr, double *ekin2, const double *pair_out, const double *ewald_out,
                    double *total, ynb_stream stream);
/* ynb_verlet_kick followed by ynb_verlet_kick_drift (the end of one step and
 * the start of the next, verlet.py:354-355 then 344-347) in one launch, with
 * the arguments of both; every output is bitwise what the two calls give
 * (2*Ekin from the velocity after the first half kick) when dt is the same
 * for both steps, as it is inside a sequence.  One kernel boundary fewer per
 * step of a multi-step sequence. */
int ynb_verlet_kick_kick_drift(const ynb_batch *b, double *pos, double *vel, double *gpos,
                               const double *gpart, const double *minv, const double *masses,
                               const double *dt, double *partial, int32_t *ctr, double *ekin2,
                               const double *pair_out, const double *ewald_out, double *total,
                               const ynb_ewald *ew, const double *charges, void *nl_work,
                               size_t nl_work_bytes, const ynb_nlist *nl, ynb_stream stream);

/* Table updates of a persistent batch in one launch.  No counterpart in the
 * reference (Yaff evaluates one system at a time; this serves batches whose
 * members change in place).  Segment i (device memory, one block each)
 * copies nbytes from src + seg.src to seg.dst or, with seg.src < 0, fills
 * them with seg.fill; elements of seg.elem = 4 or 8 bytes (8: the low 4
 * bytes of fill for 4); dst, src offsets and nbytes are multiples of elem.
 * Segments must not overlap one another. */
typedef struct ynb_segment {
  uint64_t dst;     /* device address */
  int64_t src;      /* byte offset into src, or -1: fill */
  int64_t nbytes;
  uint64_t fill;    /* fill pattern (bits of one element) */
  int32_t elem;     /* 4 or 8 */
  int32_t pad_;
} ynb_segment;
int ynb_scatter(const ynb_segment *segs, int32_t nseg, const void *src, ynb_stream stream);

/* out[s][k] = sum over slots of in[s][slot][k] for k < 10, plus optional
 * extra [nsys][10] terms; used to total a force-field evaluation on device. */
int ynb_total(const ynb_batch *b, const double *pair_out, const double *ewald_out, double *total,
              ynb_stream stream);

#ifdef __cplusplus
}
#endif

#endif
