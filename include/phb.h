/*
 * phb.h -- C ABI of ph_brush_amd, the MI355X (gfx950) implementation of
 * ph_brush's Monte Carlo energy path for protonation and ion-swap moves.
 *
 * Plain C: opaque handles, pointers and sizes, int status codes.  No
 * exception crosses this boundary; on failure a function returns a non-zero
 * PHB_ERR_* code and phb_last_error() holds the message (thread-local).
 * Every call that returns a value synchronises its stream first, like the
 * reference's "return by value after a device sync" convention.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root).  The C++ layer in include/ph_brush/
 * restores the reference signatures on top of these functions.
 *
 * Data conventions (shared with the reference):
 *   - positions wrapped into [0, L); orthorhombic box; Lz already extended
 *     by the caller when the slab correction is used (MC/src/main.cu:110-117);
 *   - device types 0-based (MC/src/main.cu:33); MC type ids 1-based
 *     (MC/include/types.hpp:51-54);
 *   - CSR rows sorted ascending; grids x-fastest ((z*ny)+y)*nx+x;
 *     half-spectrum k-grids ((kz*ny)+ky)*(nx/2+1)+kx, complex interleaved.
 */
#ifndef PHB_H
#define PHB_H

#ifdef __cplusplus
extern "C" {
#endif

#define PHB_ABI_VERSION 1

enum {
    PHB_OK = 0,
    PHB_ERR_ARG = 1,          /* invalid argument / size / range              */
    PHB_ERR_HIP = 2,          /* HIP runtime error                            */
    PHB_ERR_ALLOC = 3,        /* device or host allocation failed             */
    PHB_ERR_STATE = 4,        /* infeasible plan, no valid move, bad state    */
    PHB_ERR_UNSUPPORTED = 5,  /* option not supported by this build          */
    PHB_ERR_INTERNAL = 6
};

enum { PHB_FP64 = 0, PHB_FP32 = 1 };
enum { PHB_COUL_CUT = 0, PHB_COUL_PME = 1, PHB_COUL_UNKNOWN = 2 /* params.in value not recognised */ };

const char *phb_last_error(void);
int phb_abi_version(void);
/* Number of visible GPUs (0 when none); never fails on a CPU-only host. */
int phb_device_count(int *n);

typedef struct { double Lx, Ly, Lz; } phb_box;

/* ------------------------------------------------------------------ atoms */
/* Device-resident SoA atoms.  Replaces DeviceAtoms (MC/include/device_data.hpp:5-12)
 * and upload_atoms_to_device (MC/src/main.cu:18-41). */
typedef struct phb_atoms_s *phb_atoms;
int phb_atoms_create(int natoms, const double *x, const double *y, const double *z,
                     const double *q, const int *type0, phb_atoms *out);
int phb_atoms_destroy(phb_atoms a);
int phb_atoms_natoms(phb_atoms a, int *natoms);
/* Batched identity update.  Replaces set_atom_identity/write_device_value
 * (MC/src/mc.cu:19-41): one transfer for n atoms instead of 2n memcpys. */
int phb_atoms_set_identity(phb_atoms a, int n, const int *idx, const int *type0, const double *q);
int phb_atoms_get_identity(phb_atoms a, double *q, int *type0);
/* New (wrapped) positions for every atom; no counterpart in the reference,
 * whose positions never change after the upload (MC/src/main.cu:18-41).
 * Position-dependent state built from the old positions (neighbour lists,
 * MC objects) is not updated: rebuild it, or use a skin list. */
int phb_atoms_set_positions(phb_atoms a, const double *x, const double *y, const double *z);

/* -------------------------------------------------------------------- CSR */
/* Device CSR (rows+1 heads, nnz column ids).  Replaces DeviceCSR
 * (MC/include/device_data.hpp:14-19) and copy_csr_to_device (MC/include/neighbor.hpp:6). */
typedef struct phb_csr_s *phb_csr;
int phb_csr_create(int rows, const int *head, const int *list, phb_csr *out);
int phb_csr_destroy(phb_csr c);
int phb_csr_shape(phb_csr c, int *rows, long long *nnz);
int phb_csr_download(phb_csr c, int *head, int *list);

/* Half neighbour list: pairs i<j with min-image r^2 <= rc^2, 1-2 bonded pairs
 * (bonds12, may be NULL) excluded, rows ascending.  Replaces
 * build_half_neighbors_gpu (MC/include/neighbor.hpp:10-13, MC/src/neighbor.cu:74-125);
 * built from binned cells with wavefront-ballot compaction, bit-exact with
 * the reference's O(N^2) scan. */
int phb_nlist_build_half(phb_atoms a, const phb_box *box, phb_csr bonds12, double rc, phb_csr *out);

/* Skin (Verlet) list.  No counterpart in the reference, which rebuilds its
 * O(N^2) list from scratch (build_half_neighbors_gpu, MC/src/neighbor.cu:74-125).
 * create: the half list at rc + skin from the current positions, which are
 * snapshotted.  update: the largest min-image displacement since the snapshot;
 * when it exceeds skin / 2 the outer list and the snapshot are rebuilt
 * (*rebuilt = 1).  half: the half list at rc (the predicate of
 * phb_nlist_build_half, rows ascending) filtered from the outer list; while no
 * atom has moved more than skin / 2 since the last build it equals
 * phb_nlist_build_half(rc) at the current positions. */
typedef struct phb_verlet_s *phb_verlet;
int phb_verlet_create(phb_atoms a, const phb_box *box, phb_csr bonds12, double rc, double skin, phb_verlet *out);
int phb_verlet_destroy(phb_verlet v);
int phb_verlet_update(phb_verlet v, phb_atoms a, int *rebuilt, double *max_disp);
int phb_verlet_half(phb_verlet v, phb_atoms a, phb_csr *out);
int phb_verlet_stats(phb_verlet v, long long *builds, long long *outer_nnz, double *last_disp);

/* ----------------------------------------------------------------- params */
/* Replaces DeviceParams (MC/include/device_data.hpp:21-36) and
 * upload_params_to_device (MC/src/main.cu:43-66).  Fields marked EXT extend the
 * reference; their zero value selects the reference behaviour. */
typedef struct {
    int ntypes;
    const double *lj_eps;      /* ntypes*ntypes row-major (ti*T+tj), 0-based types */
    const double *lj_sig;
    int lj_shift;              /* shift each LJ pair to zero at the LJ cutoff */
    double rc;                 /* cutoff_nm: neighbour list and real-space cutoff */
    double lj_rc;              /* EXT: LJ cutoff (<= rc); 0 = rc */
    double epsilon_r;
    int coulombtype;           /* PHB_COUL_CUT or PHB_COUL_PME */
    double ewald_alpha;        /* > 0: used as is; else derived from ewald_rtol */
    double ewald_rtol;
    double pme_spacing;        /* grid from spacing when nx/ny/nz are 0 */
    int pme_order;             /* 4 = reference cubic; EXT: 3..8 */
    int nx, ny, nz;            /* EXT: explicit grid */
    int pme_3dc;               /* Yeh-Berkowitz slab correction */
} phb_params_desc;
typedef struct phb_params_s *phb_params;
int phb_params_create(const phb_params_desc *d, phb_params *out);
int phb_params_destroy(phb_params p);

/* Short-range energies over a half list.  Replace compute_lj_sr_energy_gpu and
 * compute_coul_sr_energy_gpu (MC/include/energy.hpp:7-18). */
int phb_energy_lj(phb_atoms a, phb_params p, phb_csr neigh, const phb_box *box, double *E);
int phb_energy_coul_cut(phb_atoms a, phb_params p, phb_csr neigh, const phb_box *box, double *E);

/* -------------------------------------------------------------------- PME */
/* PMEEnergyComponents (MC/include/pme.hpp:7-12); qcorr_term = qcorr + slab as
 * in the reference (MC/src/pme.cu:453), slab_term reports the slab part alone. */
typedef struct {
    double real_space, recip_space, self_term, qcorr_term, slab_term;
} phb_pme_components;
typedef struct phb_pme_s *phb_pme;
/* Replaces pmeCreate / pmeDestroy (MC/include/pme.hpp:42-43). */
int phb_pme_create(phb_params p, const phb_box *box, double rc, int precision, phb_pme *out);
int phb_pme_destroy(phb_pme pme);
/* The PMEPlan fields main reads (MC/src/main.cu:134-137). */
int phb_pme_info(phb_pme pme, int *nx, int *ny, int *nz, int *order, double *alpha);
/* The xy-plane transform: one workgroup per plane where the plane fits in LDS
 * (default), else row and column passes through global memory (planes of the
 * reference's pme_spacing on boxes wider than ~11 nm).  mode 1 selects the passes
 * for any plan (same results up to rounding), 0 the one-workgroup form where it
 * fits, -1 only reports; *split (may be NULL) gets the form in use.  Set it
 * before MC objects are created on the plan (their refresh gates are sized by it). */
int phb_pme_plane_mode(phb_pme pme, int mode, int *split);
/* Replaces pmeEnergy (MC/include/pme.hpp:45-49). */
int phb_pme_energy(phb_pme pme, phb_atoms a, phb_csr neigh, const phb_box *box,
                   phb_pme_components *comps, double *E);
/* Replaces the static total_energy (MC/src/mc.cu:43-60): E_LJ + E_Coulomb.
 * comps (may be NULL): [lj, real_or_cut, recip, self, qcorr, slab]. */
int phb_total_energy(phb_atoms a, phb_params p, phb_csr neigh, phb_pme pme, const phb_box *box,
                     double *comps, double *E);

/* No counterpart in the reference: the same computation as phb_total_energy for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its partial sums and
 * results live in a fixed device array of the library (8,208 doubles, 64 KiB,
 * one copy per GPU, held until the process ends) instead of per-call buffers;
 * the PME plan's grids are the plan's own.  Limit: plans whose z pass needs more
 * than 8,192 partial sums are refused (PHB_ERR_ARG).  Calls on one GPU share that
 * array: calls that overlap on one GPU (several host threads) race. */
int phb_total_energy_noalloc(phb_atoms a, phb_params p, phb_csr neigh, phb_pme pme, const phb_box *box,
                             double *comps, double *E);

/* Stage-level access for parity tests and tooling (no reference counterpart;
 * the reference keeps these buffers inside PMEPlan). Host buffers in/out. */
int phb_pme_spread(phb_pme pme, phb_atoms a, double *grid);                 /* nx*ny*nz */
int phb_pme_fft_r2c(phb_pme pme, const double *grid, double *kgrid);        /* 2*(nx/2+1)*ny*nz */
int phb_pme_influence(phb_pme pme, double *eta);                            /* (nx/2+1)*ny*nz */
/* phi = G*grid (unnormalised inverse of eta * FFT(grid)) and E = 0.5 grid.phi */
int phb_pme_potential(phb_pme pme, const double *grid, double *phi, double *E);

/* --------------------------------------------------------------------- MC */
/* Protonation + ion-swap Monte Carlo kept entirely on the device.  One trial
 * is the reference's GC move (MC/src/mc.cu:379-503): direction drawn with
 * probability a*c/(a*c+b*d), a site and an ion drawn from the pools with
 * uniform_int_distribution, Metropolis-Hastings with logH = ln S_before -
 * ln S_after, pools updated by erase(find)+push_back.  The RNG is a device
 * MT19937 with libstdc++'s Lemire and generate_canonical<double,53>, so the
 * accept/reject sequence equals the reference's under the same seed (except
 * for moves within rounding of the Metropolis threshold).  Delta-E is exact
 * (FP64): neighbour rows for LJ and real space, the reciprocal term from the
 * quadratic form 0.5 Q^T G Q through a maintained potential grid, and O(1)
 * self / net-charge / slab updates.  mu_eff is held fixed (titration at a
 * given pH: mu_eff = kT ln10 (pKa - pH)). */
typedef struct {
    int nh_n_type, nh_p_type, w_type, cl_type;   /* 1-based type ids */
    double beta;
    double mu_eff;
    double w_z_extra;          /* pool z filter: z < max site z + w_z_extra (MC/src/mc.cu:93-100) */
    int precision;             /* PHB_FP64 */
    int pending_max;           /* accepted atom changes buffered before a potential refresh (2..64; 0 = 32) */
    int recompute_interval;    /* full energy recompute (drift reset) every N trials; 0 = never */
    long long record_cap;      /* per-trial records kept (0 = none) */
    int spec_window;           /* trials evaluated concurrently per round: 1, 2, 4, 8, 16 (0 = 16) */
    int refresh;               /* potential-grid refresh: PHB_REFRESH_AUTO, _DIRECT or _KSPACE */
} phb_mc_desc;
/* Refresh of the potential grid by the pending accepted charge changes.
 * DIRECT: their stencils convolved with the real-space mesh kernel in one pass
 * over the grid (x extent up to 64 mesh points); KSPACE: structure factor of the
 * changes times the influence function, inverse FFT onto the grid; AUTO: DIRECT
 * where the grid allows it and pending_max <= 16, else KSPACE.  The two agree
 * up to rounding (the same circular convolution). */
#define PHB_REFRESH_AUTO 0
#define PHB_REFRESH_DIRECT 1
#define PHB_REFRESH_KSPACE 2
typedef struct {
    long long attempted, accepted, accepted_forward, accepted_backward;
    int nprot, nsites, status;         /* status 1: no valid pair to move (MC/src/mc.cu:398-402) */
    double E;                          /* running total energy */
    double E_recomputed;               /* last full recompute (NaN if none) */
    double max_drift;                  /* max |running - recomputed| over recomputes */
    long long refreshes, recomputes, launches;
    /* with profiling on: summed HIP-event times on the library stream */
    double chain_ms, refresh_ms, recompute_ms;
} phb_mc_stats;
typedef struct phb_mc_s *phb_mc;
/* Builds pools (build_plan, MC/src/mc.cu:67-122), full neighbour rows with
 * real-space coefficients, spline stencils, the mesh kernel and the potential
 * grid; seeds the device MT19937 like std::mt19937(seed).  The atoms handle is
 * updated in place by every accepted move and must outlive the MC object.
 * coulombtype pme: pme is the plan (FP64).  coulombtype coul_cut
 * (total_energy's other branch, MC/src/mc.cu:52-56): pme may be NULL and is not
 * used; the chain runs on a plan of its own with alpha 0 (pair coefficient
 * elF/r), no reciprocal part and no slab term. */
int phb_mc_create(phb_atoms a, phb_params p, phb_csr neigh, phb_pme pme, const phb_box *box,
                  const phb_mc_desc *d, unsigned seed, phb_mc *out);
int phb_mc_destroy(phb_mc mc);
/* Runs ntrials more trials (replaces the trial loop of phase_gc_mu_control at fixed mu). */
int phb_mc_run(phb_mc mc, long long ntrials, phb_mc_stats *out);
int phb_mc_stats_get(phb_mc mc, phb_mc_stats *out);
/* First n per-trial records (any output may be NULL); u is NaN when no uniform was drawn. */
int phb_mc_record(phb_mc mc, long long n, signed char *fwd, int *i, int *j, signed char *acc, double *dE,
                  double *u);
/* Delta-E of n independent candidate moves (site i, ion j, fwd 1 = protonate)
 * against the current state, without accepting them. */
int phb_mc_delta_e(phb_mc mc, int n, const int *i, const int *j, const int *fwd, double *dE);
/* Full recompute now (drift reset); returns the recomputed total energy. */
int phb_mc_recompute(phb_mc mc, double *E);
/* MT19937 state: 624 words then the index (625 unsigned), the stream form of
 * std::mt19937 (operator<< / operator>>).  get: the engine after the last run;
 * set: the next run continues from that engine, so a caller's std::mt19937 is
 * handed in and out as the reference passes `std::mt19937& rng` through
 * phase_gc_mu_control and phase2_swap_jitter (MC/include/mc.hpp:58-81,
 * MC/src/main.cu:180-206).  set rejects an index above 624. */
int phb_mc_rng_state(phb_mc mc, unsigned *state);
int phb_mc_set_rng_state(phb_mc mc, const unsigned *state);
/* Sets the fixed mu_eff of the chain's next runs (plain and delayed-acceptance):
 * a titration moves one chain from pH to pH, as phase_gc_mu_control moves
 * Params::mu_eff between blocks (MC/src/mc.cu:349, :530).  A batch built from
 * the chain keeps the mu_eff it was built with. */
int phb_mc_set_mu(phb_mc mc, double mu_eff);
/* Tile shape of the chain's direct potential refresh: ty rows along y (8 or 4) by tz planes along
 * z (4 or 2); default 8 x 4.  Pending changes are folded in first.  The refresh has no counterpart
 * in the reference, which recomputes the potential per trial. */
int phb_mc_set_refresh_tile(phb_mc mc, int ty, int tz);
/* std::mt19937(seed) in the stream form above (host only). */
int phb_rng_seed_state(unsigned seed, unsigned *state);
/* The chains' device engine on its own (no reference counterpart; it checks the
 * device form of std::mt19937, uniform_int_distribution<int> and
 * uniform_real_distribution<double>, MC/src/mc.cu:62-65, :368-369, draw by
 * draw).  From an engine state (625 words), draw k is
 * uniform_int_distribution<int>(0, hi[k]) when hi[k] >= 0, else a uniform double
 * in [0, 1) (out[k] = its IEEE bits).  Draws go in rounds of 64 read at fixed
 * stream positions, a round with a Lemire rejection redrawn serially, as the
 * chains propose.  state_after (may be NULL): the engine afterwards;
 * serial_rounds (may be NULL): rounds that took the serial path. */
int phb_rng_draws(const unsigned *state, long long n, const int *hi, unsigned long long *out,
                  unsigned *state_after, long long *serial_rounds);

/* What the last phb_phase1_run did, block by block: the events its reference
 * counterpart reports as it runs (MC/src/mc.cu:153-155, :356-358, :399, :471-479,
 * :508-516, :534-545), so that a caller can write the same report lines.
 * Counts are of that run; mu_over_kT is beta * mu_eff at the event (after the
 * update for PHB_P1_MU_UPDATE); E the chain energy at the event. */
enum {
    PHB_P1_DIRECT = 1,      /* direct flips applied: attempted = flips */
    PHB_P1_AT_TARGET = 2,   /* already at the target: no trial made */
    PHB_P1_CHUNK = 3,       /* a block of gc_block trials ended short of the target */
    PHB_P1_MU_UPDATE = 4,   /* an epoch ended: mu -= kT eta logit_diff w_info */
    PHB_P1_REACHED = 5,     /* the target was reached at trial step_in_block of a block */
    PHB_P1_NO_PAIRS = 6     /* no valid pair to flip: the run stopped */
};
typedef struct {
    int kind;
    int epoch;                          /* 0-based epoch */
    long long attempted, accepted;      /* of the run so far */
    long long fwd_total, bwd_total;     /* accepted forward / backward moves of the run */
    long long acc_since, att_since;     /* accepted / attempted since the last mu update */
    long long block_attempted, block_accepted;
    int step_in_block;                  /* PHB_P1_REACHED: trials of the block up to the target */
    int early_stop;                     /* PHB_P1_MU_UPDATE: the epoch ended on epoch_attempt_max */
    int nprot, nh, target_count, remaining, r0, nacc_target, amax, gc_block;
    double mu_over_kT, E, f, w_info;
} phb_phase1_event;
/* Phase 1 as main() runs it (MC/src/main.cu:180-197) on an MC object: build_plan's
 * checks (MC/src/mc.cu:67-122; errors as PHB_ERR_STATE with the reference's
 * message), then apply_direct_protonation (MC/src/mc.cu:124-157) when target_pct
 * > direct_pct_threshold, else phase_gc_mu_control (MC/src/mc.cu:320-546) on the
 * device chain: mu from the target fraction (clamped to +-10 kT), blocks of
 * gc_block trials, an epoch ends between blocks after Nacc_target = ceil(nacc_frac
 * * remaining) acceptances or epoch_attempt_max trials, then mu -= kT eta
 * logit_diff w_info; the run stops at the trial whose acceptance reaches the
 * target.  The pools handed to phase 2 are kept on the object (empty unless the
 * target was reached mid-run or the direct path ran, as in the reference).
 * max_trials bounds the run (0 = none; the reference's loop has no bound). */
typedef struct {
    double target_pct;         /* target_protonation_pct */
    int direct_pct_threshold;  /* reference default 90 */
    int gc_block;              /* gc_block_fixed (0 = 2000) */
    int epoch_attempt_max;     /* sgmc_epoch_attempt_max (0 = 100000) */
    double nacc_frac;          /* sgmc_nacc_frac_round (<= 0: 0.10) */
    double mu_eta;             /* (<= 0: 0.3) */
    long long max_trials;
    /* optional: called on the host with each event as it happens (the events of
     * phb_phase1_events, in order), e.g. to write report lines while the run goes on */
    void (*on_event)(const phb_phase1_event *ev, void *user);
    void *user;
} phb_phase1_desc;
typedef struct {
    long long attempted, accepted, accepted_forward, accepted_backward;
    int blocks_done, Nh, Nprot, target_count, reached_target, direct, status;
    double E, mu_eff;
} phb_phase1_result;
int phb_phase1_run(phb_mc mc, const phb_phase1_desc *d, phb_phase1_result *out);
/* The lists phase 1 handed to phase 2 (Phase1Result's nh_n_remaining, nh_p_mc,
 * w_remaining, cl_mc, MC/include/mc.hpp:19-27), in that order: sizes[4] always;
 * lists (capacity cap ints) when non-null, concatenated.  All empty before a
 * phase-1 run or when it ended without reaching its target. */
int phb_phase1_handover(phb_mc mc, int sizes[4], int *lists, long long cap);
/* n = events recorded (always); ev filled with the first min(n, cap) when non-null. */
int phb_phase1_events(phb_mc mc, phb_phase1_event *ev, int cap, int *n);

/* Phase 2: the 4-point and W-Cl-only move rule of phase2_swap_jitter
 * (MC/src/mc.cu:159-315) on an MC object, continuing its engine stream.  begin:
 * the pools phb_phase1_run handed over, else rebuilt in ascending atom order from
 * the current identities as the phase-1 loop hands them over when it reaches its
 * target (MC/src/mc.cu:480-494); W and Cl below the pool z cut; the mode chosen as the reference does (MC/src/mc.cu:193-207): the
 * 4-point move unless NH_N is empty or no 4-point move exists, then W-Cl only,
 * else no valid proposals.  run: up to max_attempts more trials; a trial is
 * accepted if dE <= 0 or exp(clamp(-beta dE, +-40)) > u (u drawn only when
 * dE > 0); stops after `patience` trials without E < best_E - best_eps. */
typedef struct {
    int patience;
    double best_eps;           /* reference: 1e-9 */
    long long record_cap;      /* per-trial records kept */
} phb_phase2_desc;
typedef struct {
    long long attempted, accepted;
    int mode;                  /* 4: 4-point, 2: W-Cl only, 0: no valid proposals */
    int stop;                  /* 0 running, 1 patience reached, 2 no valid proposals */
    int since_best, status;
    double E, best_E;
} phb_phase2_stats;
/* The pools phase 2 starts from, given by the caller in phb_phase1_handover's
 * order (nh_n, nh_p, w, cl; sizes[4], lists concatenated): the Phase1Result
 * lists phase2_swap_jitter receives (MC/include/mc.hpp:61-70, MC/src/mc.cu:172-184).
 * They replace any hand-over of phb_phase1_run; begin then filters W and Cl by
 * the pool z cut.  Indices are checked against the atom count. */
int phb_phase2_set_pools(phb_mc mc, const int sizes[4], const int *lists);
int phb_phase2_begin(phb_mc mc, const phb_phase2_desc *d, phb_phase2_stats *out);
int phb_phase2_run(phb_mc mc, long long max_attempts, phb_phase2_stats *out);
/* First n records: atoms (4 per trial in draw order iP, iN, jCl, jW; W-Cl only: jCl, jW, -1, -1),
 * accepted flag and exact Delta-E.  Records are kept for trials first .. first + record_cap - 1
 * of the phase (first = 0 after begin); record_from moves that window for the trials still to
 * run, so that a long phase 2 is recorded window by window. */
int phb_phase2_record(phb_mc mc, long long n, int *atoms, signed char *acc, double *dE);
int phb_phase2_record_from(phb_mc mc, long long first);

/* Time every chain launch, refresh and recompute with HIP events: on = 1 also runs the
 * delayed-acceptance chain's in-kernel phase clocks (phb_da_stats.phase_us), on = 2 the
 * events alone (the kernels run as unprofiled), 0 off. */
int phb_mc_set_profiling(phb_mc mc, int on);

/* ------------------------------------------------------------ file formats
 * The reference's on-disk inputs and outputs, host code (MC/include/io.hpp,
 * MC/src/io.cpp).  A file the reference reads gives the same values here, and a
 * file written here is byte-identical to the reference's from the same system.
 * Errors carry the reference's messages (PHB_ERR_ARG). */
typedef struct phb_data_s *phb_data;
typedef struct {
    int natoms, nbonds, ntypes;     /* header counts ("N atoms", "N bonds", "N atom types") */
    int nbond_rows;                 /* Bonds rows kept: both atom ids known */
    phb_box box;                    /* xhi-xlo, yhi-ylo, zhi-zlo */
} phb_data_info;
/* read_lammps_full (MC/include/io.hpp:7, MC/src/io.cpp:79-146): the header
 * before "Atoms", the Atoms rows (id mol type q x y z, file order, types
 * 1-based) and the Bonds rows (atom ids mapped to 0-based row indices). */
int phb_data_read(const char *path, phb_data *out, phb_data_info *info);
/* Copies the rows out: natoms entries per atom column, nbond_rows per bond
 * column; a NULL column is skipped. */
int phb_data_get(phb_data d, int *id, int *mol, int *type, double *q, double *x, double *y, double *z,
                 int *bond_id, int *bond_type, int *bond_i, int *bond_j);
int phb_data_destroy(phb_data d);
typedef struct {
    int natoms, nbonds, ntypes;     /* header counts written */
    int natoms_rows, nbond_rows;    /* rows written */
    const int *id, *mol, *type;     /* 1-based types */
    const double *q, *x, *y, *z;
    const int *bond_id, *bond_type, *bond_i, *bond_j;   /* bond_i/j: 0-based rows of the atom columns */
} phb_data_view;
/* write_lammps_full (MC/include/io.hpp:17, MC/src/io.cpp:326-368): box from 0 to
 * box->L per axis (the data-file box, not one extended for the slab correction). */
int phb_data_write(const char *path, const phb_data_view *v, const phb_box *box);
/* wrap_positions (MC/include/io.hpp:14, MC/src/io.cpp:303-316): into [0, L) per axis. */
int phb_wrap_positions(int n, double *x, double *y, double *z, const phb_box *box);
/* Ewald alpha from a real-space tolerance as read_params derives it
 * (MC/src/io.cpp:46-77): erfc^-1(clamp(rtol, 1e-12, 0.5)) / rc, rtol <= 0 read as 1e-5. */
int phb_ewald_alpha_from_rtol(double rc, double rtol, double *alpha);

/* params.in (read_params, MC/include/io.hpp:11, MC/src/io.cpp:171-301): "key =
 * value" lines ('-' in a key read as '_', unknown keys ignored, '#' starts a
 * comment line) and "lj ti tj eps sig" lines; then the reference's defaults,
 * alpha from ewald_rtol when PME has no ewald_alpha, the LJ(rc) shift table and
 * its range checks.  Two keys are this library's: lj_cutoff_nm (a separate LJ
 * cutoff, 0: rc) and pme_grid = nx ny nz (0: from pme_spacing). */
typedef struct {
    double rc;
    int ntypes;
    int lj_shift;
    int coulombtype;                /* PHB_COUL_CUT, PHB_COUL_PME or PHB_COUL_UNKNOWN */
    double epsilon_r, ewald_rtol, ewald_alpha, pme_spacing;
    int pme_order, pme_3dc;
    double pme_3dc_zfac;
    int nh_n_type, nh_p_type, w_type, cl_type;   /* 1-based */
    double target_protonation_pct;
    int rng_seed, energy_interval, patience, max_attempts;
    double beta, w_z_extra_nm;
    int direct_pct_threshold;
    double mu_eta;
    int gc_block_fixed, sgmc_epoch_attempt_max;
    double sgmc_nacc_frac_round;
    int phase2_wc;
    double lj_rc;                   /* extension: lj_cutoff_nm */
    int nx, ny, nz;                 /* extension: pme_grid */
} phb_params_file;
/* lj_eps, lj_sig, lj_shift_tbl: ntypes*ntypes row-major (ti-1)*ntypes+(tj-1), or NULL. */
int phb_params_read(const char *path, int ntypes, phb_params_file *out, double *lj_eps, double *lj_sig,
                    double *lj_shift_tbl);

/* ----------------------------------------------------------- generator */
/* Synthetic brush (stands in for the missing LAMMPS inputs). */
typedef struct {
    int chains_x, chains_y;    /* graft lattice */
    double graft_spacing;      /* first graft point at spacing/2 */
    double graft_z;            /* z of each chain's first bead */
    int chain_len;
    int n_perm;                /* beads 0..n_perm-1 of each chain are permanently +1 */
    double p_prot;             /* initial protonation probability of titratable beads */
    int n_salt;                /* salt pairs (cation + Cl-) */
    double Lx, Ly, Lz;         /* data-file box */
    double bond;               /* random-walk step */
    double min_dist;           /* overlap rejection distance */
    unsigned long long seed;
    int type_perm, type_cation, type_nh_n, type_nh_p, type_w, type_cl;   /* 1-based */
} phb_brush_spec;
/* Call with x == NULL to query natoms/nbonds; types written 1-based. */
int phb_gen_brush(const phb_brush_spec *spec, int cap, double *x, double *y, double *z, double *q,
                  int *type1, int *mol, int *bond_i, int *bond_j, int bond_cap, int *natoms, int *nbonds);


/* ------------------------------------------------- learned surrogate + delayed acceptance
 * No counterpart in the reference (it has no learned part).  A move's
 * descriptor is d = h_i - h_j: 48 counts per centre (12 shells of r_desc/12 x
 * species 0 charged bead, 1 neutral titratable bead, 2 cation, 3 anion), each
 * centre excluding the other.  The surrogate is dE* = B + s f(d): B the exact
 * reciprocal-space potential term sum_a dq_a W_a . phi, the mesh self terms, a
 * pair term and the O(1) corrections; f the network (48 -> 128 -> 128 -> 1,
 * GELU, BF16 MFMA with FP32 accumulation) trained on t = s (dE - B), s = +1
 * forward.
 * Flat FP32 parameter layout: W1[48][128], b1[128], W2[128][128], b2[128],
 * w3[128], b3 (PHB_NN_TRAIN trainable values), then in_scale[48], out_mean,
 * out_std. */
#define PHB_NN_IN 48
#define PHB_NN_HIDDEN 128
#define PHB_NN_TRAIN 22913
#define PHB_NN_PARAMS 22963
typedef struct phb_surrogate_s *phb_surrogate;
/* Network with He-uniform initial weights (std::mt19937_64(seed)) and a device
 * dataset of `capacity` (descriptor, target) samples. */
int phb_surrogate_create(long long capacity, unsigned seed, phb_surrogate *out);
int phb_surrogate_destroy(phb_surrogate s);
int phb_surrogate_set_params(phb_surrogate s, const float *theta);    /* PHB_NN_PARAMS; resets Adam */
int phb_surrogate_get_params(phb_surrogate s, float *theta);
/* y = out_mean + out_std * f(d) for n raw descriptors (n x 48). */
int phb_surrogate_forward(phb_surrogate s, int n, const short *d, float *y);
/* Gradient of mean((f_norm(d) - (t - out_mean)/out_std)^2) over n rows (n % 64 == 0), no update. */
int phb_surrogate_grad(phb_surrogate s, int n, const short *d, const float *t, float *grad, double *loss);
int phb_surrogate_data_add(phb_surrogate s, int n, const short *d, const float *t);
int phb_surrogate_data_get(phb_surrogate s, long long first, int n, short *d, float *t);
int phb_surrogate_data_size(phb_surrogate s, long long *n);
int phb_surrogate_data_clear(phb_surrogate s);
typedef struct {
    int epochs;
    int batch;                 /* multiple of 64; 0 = 4096 */
    double lr;                 /* Adam step size */
    unsigned seed;             /* shuffling (std::mt19937_64) */
    long long n_train;         /* first n_train samples are trained on; 0 = all */
    int normalize;             /* set in_scale / out_mean / out_std from the training data first */
} phb_train_desc;
typedef struct {
    long long steps;
    double first_loss, last_loss;   /* normalised mean-squared loss of the first and last step */
    double ms;
} phb_train_stats;
/* Adam training on the GPU (own forward/backward kernels). */
int phb_surrogate_train(phb_surrogate s, const phb_train_desc *d, phb_train_stats *st);
int phb_surrogate_loss_history(phb_surrogate s, long long n, float *loss);
/* Error of out_mean + out_std f(d) against the stored targets over a dataset range (target units). */
int phb_surrogate_eval(phb_surrogate s, long long first, long long n, double *rmse, double *mae);

typedef struct {
    const signed char *species_of_type;  /* ntypes entries (0-based types): -1 not counted, else 0..3 */
    int ntypes;
    double r_desc;                       /* histogram radius */
    int candidates;                      /* surrogate screens per round, 1..64 (0 = 64) */
    int real_space;                      /* 1: the surrogate's physics part includes the real-space potential
                                            of the listed neighbours of the site and the ion (exact, kept in
                                            fixed point); 0: it does not.  A network is trained with the
                                            physics part of the chain it screens for. */
    int local_pairs;                     /* 1: local-pair form.  The physics part takes the potential at the site
                                            and the ion as far_a + q_a s_a + T_a, T_a = sum over the pool atoms b
                                            of L(a) of q_b C_ab plus, for the pool atoms outside L(a), their
                                            charge changes since phb_da_freeze times a smoothed continuum
                                            reciprocal pair kernel; T is kept exactly in fixed point as moves are
                                            accepted and far_a (everything else) is frozen at the state
                                            phb_da_freeze saw.  L(a) holds the pool atoms within r_local of a and
                                            those in its lateral column (|dx|, |dy| <= 5 mesh spacings, where the
                                            order-P mesh couples z-aligned atoms), C_ab = W_a^T G W_b + k_ab.
                                            The surrogate then needs no potential refresh after an accepted move,
                                            and a launch holds up to pending_max/2 of them (the exact stage takes
                                            the pending changes as the plain chain does).  real_space is ignored.
                                            0: the potential term is W_a . phi of the current state, one accepted
                                            move per launch. */
    double r_local;                      /* local-pair radius: 0 = r_desc; rc <= r_local <= r_desc */
} phb_da_desc;
/* Builds the r_desc neighbour rows and per-atom histograms of a chain (and, in the
 * local-pair form, the local pair rows and their coefficients). */
int phb_da_attach(phb_mc mc, const phb_da_desc *d);
/* Local-pair form: freezes far_a at the current state.  The first run, collection or
 * prediction after phb_da_attach freezes by itself; later runs keep the frozen field
 * (a chain is exact for a field fixed before it starts), so call this at the start
 * of each new chain.  PHB_ERR_STATE without the local-pair form. */
int phb_da_freeze(phb_mc mc);
/* Descriptors (n x 48) and pair distances of moves in the chain's current state. */
int phb_da_descriptors(phb_mc mc, int n, const int *i, const int *j, short *d, double *rij);
/* Appends (descriptor, s (dE - B)) of n moves, dE exact, to the dataset. */
int phb_da_collect(phb_mc mc, phb_surrogate s, int n, const int *i, const int *j, const int *fwd);
/* Surrogate dE* of n moves. */
int phb_da_predict(phb_mc mc, phb_surrogate s, int n, const int *i, const int *j, const int *fwd, double *dE_star);
typedef struct {
    long long attempted, stage1_accepted, accepted, rounds, launches;
    int nprot, nsites, status;
    double E;
    double chain_ms, refresh_ms;         /* with phb_mc_set_profiling */
    double phase_us[8];                  /* with profiling: chain time by phase (launch prologue, proposals,
                                            descriptors, MLP, stage 1, exact dE, stage 2 + updates, epilogue) */
} phb_da_stats;
/* Runs ntrials delayed-acceptance trials: stage 1 on dE*, stage 2 on the exact dE,
 * same proposals, pools and engine stream as phb_mc_run. */
int phb_da_run(phb_mc mc, phb_surrogate s, long long ntrials, phb_da_stats *out);
/* Per-trial record: stage 0 rejected at stage 1, 1 accepted, 2 rejected at stage 2; dE* of the move and of its
 * reverse from the proposed state (NaN when stage 1 rejected); dE NaN when not evaluated. */
int phb_da_record(phb_mc mc, long long n, signed char *fwd, int *i, int *j, signed char *stage, double *dE_star,
                  double *dE_star_rev, double *dE);


/* ------------------------------------------------------------ batched replicas
 * No counterpart in the reference (one system per process there).  Replicas of
 * a titration grid (own system, PME plan and chain each; same spline order,
 * FP64) advance together: one workgroup per replica in each chain launch, the
 * potential refreshes of all replicas in two launches, the (chain, refresh)
 * launch pairs captured in a hipGraph; or, in the persistent form (every replica
 * on the direct refresh), one cooperative launch per sweep in which each
 * replica's workgroup runs its chain to the sweep target and the potential
 * refreshes are folded tile by tile by that workgroup and the launch's other
 * workgroups.  Each replica's accept/reject sequence is the one phb_mc_run would
 * give it. */
typedef struct phb_batch_s *phb_batch;
typedef struct {
    int graph_pairs;           /* (chain, refresh) launch pairs per replay; 0 = 4 */
    int use_graph;             /* capture the pairs in a hipGraph (1) or launch them directly (0) */
    int spec_window;           /* out: current speculation window */
    int pending_max;           /* out: current pending-change budget */
    long long launch_cap;      /* out: trials after which a chain launch stops per replica (0: none) */
    int persistent;            /* in: start in the persistent form; out: the current form */
} phb_batch_desc;
typedef struct {
    long long attempted, accepted;   /* totals over the replicas (since creation) */
    long long launches, replays;     /* chain launches and replays of this call */
    double ms, trials_per_s;         /* wall time of this call; autotune: rate of the chosen shape */
    int spec_window, pending_max, graph_pairs;
    long long launch_cap;            /* a chain launch waits for its slowest replica: the trial cap bounds it */
    int persistent;                  /* 1: one persistent launch per sweep */
    int refresh_ty, refresh_tz;      /* direct-refresh tile: rows along y, planes along z */
} phb_batch_stats;
int phb_batch_create(const phb_mc *mcs, int n, const phb_batch_desc *d, phb_batch *out);
int phb_batch_destroy(phb_batch b);
/* `sweeps` sweeps: every replica attempts as many trials as it has titratable sites per sweep. */
int phb_batch_run(phb_batch b, int sweeps, phb_batch_stats *st);
/* After 2 x `sweeps_per_candidate` warm-up sweeps, times speculation window (4, 8, 16) x pending
 * budget (2, 4, 8, 16) x graph length (2, 4, 8) on the batch (each candidate runs `sweeps_per_candidate` real sweeps), the
 * persistent form at each window x budget where every replica qualifies, per-launch trial caps
 * on the fastest graph shape, then the two forms' fastest shapes again back to back (twice the
 * sweeps), then the direct refresh's tile shape (8 or 4 rows by 4 or 2 planes) on the faster,
 * and keeps the fastest.  Launch boundaries and tile shapes do not change any replica's sequence. */
int phb_batch_autotune(phb_batch b, int sweeps_per_candidate, phb_batch_stats *best);
/* Sets a launch shape without timing: speculation window (4, 8 or 16), pending budget (2..MC_KMAX, even),
 * (chain, refresh) pairs per replay (>= 1) and the per-launch trial cap (0: none). */
int phb_batch_tune(phb_batch b, int spec_window, int pending_max, int graph_pairs, long long launch_cap);
/* 1: the persistent form (one cooperative launch per sweep; PHB_ERR_UNSUPPORTED unless every replica
 * uses the direct refresh and the device holds a workgroup per replica at once), 0: graph replays.
 * A persistent sweep whose waits pass 30 s of wall clock stops with an error. */
int phb_batch_mode(phb_batch b, int persistent);
/* Direct-refresh tile shape of every replica (ty rows along y: 8 or 4; tz planes along z: 4 or 2);
 * no counterpart in the reference, whose potential is recomputed per trial. */
int phb_batch_set_refresh_tile(phb_batch b, int ty, int tz);
/* Per replica, of the last persistent sweep: wall time from the launch's start to the replica's
 * finish and the part of it its chain spent on refreshes (waiting for and folding tiles), in ms. */
int phb_batch_replica_times(phb_batch b, double *run_ms, double *refresh_ms);

#ifdef __cplusplus
}
#endif
#endif /* PHB_H */
