/*
 * pybrown_amd.h -- C ABI of libpybrown_amd.so, the MI355X (gfx950) engine for
 * pyBrown's hydrodynamic step.
 *
 * Reference citations are relative to the pyBrown source tree (src/bdsim/).
 * Matrices crossing this boundary are dense, row-major, 3N x 3N doubles (the
 * layout pyBrown obtains with numpy.reshape(..., (3N, 3N)),
 * pyBrown/diffusion.py:54).  All pointers are host pointers unless a name
 * ends in _dev.
 */
#ifndef PYBROWN_AMD_H
#define PYBROWN_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * 1. Drop-in replacements for libpyBrown.so.  Same symbol names, argument
 *    order, meaning and (absence of) return codes as the reference; pyBrown
 *    loads them with ctypes.cdll.LoadLibrary (pyBrown/diffusion.py:24-26,
 *    pyBrown/bead.py:24-26).  On a HIP failure they print a message to
 *    stderr and leave it in pbx_last_error(); the Python layer raises.
 * ------------------------------------------------------------------------ */

/* clib/diff_tensor.c:73  -- free-space RPY mobility (overlapping and contained
 * beads included).  `pointers` is the packed pair list -rij[p][q], p<q, of
 * pyBrown/diffusion.py:40.  Writes all 9N^2 entries of M_matrix. */
void RPY_M_matrix(double* as, double* pointers, int number_of_beads, double* M_matrix);

/* clib/diff_tensor.c:170 -- periodic RPY by Smith-Ewald summation; real-space
 * images with |m|_1 <= m_max, reciprocal vectors with 0 < |n|_1 <= n_max
 * (n_max <= 8); alpha in box units. */
void RPY_Smith_M_matrix(double* as, double* pointers, double box_length, double alpha, int m_max, int n_max,
                        int number_of_beads, double* M_matrix);

/* clib/diff_tensor.c:270 -- Jeffrey-Onishi two-sphere resistance, 18 packed
 * values (blocks 00, 11, 10 as xx,yy,zz,xy,xz,yz). */
void JO_2B_R_matrix(double ai, double aj, double rx, double ry, double rz, double* R_matrix);

/* clib/diff_tensor.c:346 -- lubrication correction; ACCUMULATES (+=) into
 * correction_matrix like the reference. */
void JO_R_lubrication_correction_matrix(double* as, double* pointers, int number_of_beads, double cutoff_distance,
                                        int cichocki_correction, double* correction_matrix);

/* clib/bead.c:27 -- minimum-image pointer matrix, N x N x 3; diagonal entries
 * keep the caller's values. */
void pointer_pbc_matrix(double* positions, int number_of_beads, double box_length, double* pointers);
/* bead.c:99 -- pointers[i][j] = wrap(r_immobile_j - r_mobile_i), shape
 * (number_of_mobile, number_of_immobile, 3); every entry written. */
void pointer_immobile_pbc_matrix(double* positions_mobile, double* positions_immobile, int number_of_mobile,
                                 int number_of_immobile, double box_length, double* pointers);

/* clib/bead.c:169 -- 1 if any non-bonded pair overlaps, else 0 (-1 on a HIP
 * error).  connection_matrix is N x N int32 (may be NULL = no bonds). */
int check_overlaps(double* positions, double* as, int number_of_beads, double box_length, double overlap_treshold,
                   int* connection_matrix);

/* ------------------------------------------------------------------------
 * 2. Dense kernels at the reference's numpy.linalg call sites.
 * ------------------------------------------------------------------------ */

/* pyBrown/box.py:554 np.linalg.cholesky -- lower factor, upper triangle set
 * to zero.  Returns 0, or k > 0 when the pivot at column k-1 is not
 * positive (numpy raises LinAlgError there), or a negative error. */
int pbx_cholesky(const double* A, int n, double* L);
/* As pbx_cholesky, with the trailing updates on FP32 matrix cores (operands
 * rounded to FP32, FP32 accumulation added into the FP64 matrix) and the
 * diagonal tiles factored in FP64.  No reference counterpart: the mixed-
 * precision noise factor measured against the FP64 one. */
int pbx_cholesky_mixed(const double* A, int n, double* L);

/* pyBrown/box.py:514,524,537 np.linalg.inv of an SPD matrix, times scale. */
int pbx_spd_inverse_host(const double* A, int n, double scale, double* X);
/* As pbx_spd_inverse_host, with the factorisation's trailing updates and the
 * tile products of the triangular inverse and of inv(L)^T inv(L) on FP32
 * matrix cores (diagonal tiles, their inverses and all stored values FP64).
 * refine = 1 adds one Newton-Schulz step X + X (I - A X) whose residual is
 * formed in FP64.  No reference counterpart: the mixed-precision resistance
 * inversion measured against the FP64 one. */
int pbx_spd_inverse_mixed(const double* A, int n, double scale, int refine, double* X);

/* x = inv(A) b for the n x n row-major SPD A, through A = L L^T and two
 * chained triangular solves (y = inv(L) b is returned too).  Replaces the
 * product of an explicit inverse with a vector where box.py:303-308 reads
 * D_mid = kBT inv(R_mid) (box.py:537) only as D_mid w.  The solves run reps
 * times; ms (may be null) receives their mean device time per pair.
 * Returns 0, k>0 (non-positive pivot at column k-1) or a negative error. */
int pbx_spd_solve(const double* A, int n, const double* b, int reps, double* y, double* x, double* ms);

/* pbx_cholesky for a matrix already on the device, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  D_dev: the n x n SPD matrix, row-major, leading
 * dimension ld_D >= n (read only).  B_dev: np = round_up(n, 64) rows,
 * leading dimension ld_B >= np (even), 16-byte aligned; receives the lower
 * factor in its leading n x n block with the upper triangle zero, the
 * identity in rows and columns n..np-1, the same bits as pbx_cholesky with
 * the default Cholesky variant.  stream: a hipStream_t (NULL = the null
 * stream); the call returns after the stream has run it.  Limit: n <= 12288
 * (larger orders return PBX_ERR_ARG).  The work area and status word live in
 * one static device array per GPU (6.3 MB, held until the process ends),
 * shared by every call on that GPU: calls that overlap on one GPU race.
 * Returns 0, k > 0 for a non-positive pivot at column k-1, or a negative
 * error. */
int pbx_cholesky_noalloc(const double* D_dev, int n, int ld_D, double* B_dev, int ld_B, void* stream);
/* Selects the Cholesky used by every factorisation in this process:
 * 0 = the library default (6);
 * 1 = one launch per panel, column-by-column diagonal tiles;
 * 2 = one launch per panel, 8-column-panel diagonal tiles;
 * 4 = one launch per panel, the solve of the next tile column folded in,
 *     wave-specialised diagonal tiles;
 * 6 = two tile columns per launch (rank-128 trailing updates).
 * Returns PBX_ERR_ARG for any other value (3 and 5 named persistent
 * single-launch variants that are no longer built). */
int pbx_set_potrf_variant(int variant);
/* The Cholesky variant in use (1, 2, 4 or 6). */
int pbx_get_potrf_variant(void);
/* Variant 6: when on, the two helper workgroups of each launch leave their
 * jobs (the panel updates of tiles (c0+1, c0) and (c0+1, c0+1)) to the
 * diagonal chain, which is the path taken when they are not yet running
 * (other launches fill the compute units).  The factor is the same to the
 * bit either way.  For tests; off by default.  No counterpart in the
 * reference (numpy.linalg.cholesky has no such choice). */
int pbx_set_potrf_chain_jobs(int on);
/* Variant 6 launch shape: orders of up to `tiles` tile columns (default
 * 64) run one workgroup per compute unit and update every trailing tile in
 * every launch (rank-128); larger ones run two per CU and update the next
 * two tile columns in most launches and the whole trailing matrix (rank
 * 128 k) every k-th, which divides the trailing matrix's HBM traffic by
 * about k.  The factors agree to rounding.  Process-wide.  No counterpart
 * in the reference. */
int pbx_set_potrf_own_cu_tiles(int tiles);
int pbx_get_potrf_own_cu_tiles(void);
/* Larger orders update each far trailing tile once every k launches (1..8,
 * default 4): up to 128 tile columns a rolling 1/k of the far tile-column
 * pairs in each launch, beyond that the whole trailing matrix in every k-th
 * launch; the next launch's two tile columns in every launch. */
int pbx_set_potrf_far_period(int k);
/* Orders of one workgroup per compute unit: 0 updates every trailing tile in
 * every launch; k = 1..8 rolls 1/k of the far tile-column pairs through each
 * launch (default 4; C2, n = 3000: 1.8 % faster).  The factors agree to
 * rounding.  Process-wide.  No counterpart in the reference. */
int pbx_set_potrf_own_cu_roll(int k);
/* Per-order launch shapes, as chosen by a tuner (pybrown_amd.autotune):
 * codes[T] for T tile columns (T < len <= 1024): 0 = the settings above,
 * 4 = variant 4, 61 = variant 6 one workgroup per CU, 62 = variant 6 two
 * per CU.  len = 0 clears the table.  The mixed-precision factor keeps
 * variant 6.  No counterpart in the reference. */
int pbx_set_potrf_table(const int* codes, int len);
/* The variant and own-CU choice a factorisation of T tile columns uses. */
int pbx_get_potrf_shape(int T, int* variant, int* own_cu);
/* Milliseconds per factorisation (the copy of D into B plus the
 * factorisation, as pbx_cholesky runs them on the device) of a synthetic
 * diagonally dominant SPD matrix of order n, with the launch shapes in force;
 * one untimed run, then the mean of `reps`.  For the tuner. */
int pbx_time_potrf(int n, int reps, int mixed, double* ms);
/* "<name> (<architecture>)" of the current device, NUL-terminated in out. */
int pbx_device_name(char* out, int len);
/* Test hook: the k-th launch sequence from now that reports through a
 * simulation's or call's status word (a factorisation, an SPD inverse's
 * factorisation, the midpoint's B0^T solve) ends with that word set as if
 * one of its bounded waits had run out, so tests can check from the host
 * that the failure is reported by the step that ran it.  0 disarms (the
 * default); every kernel still runs to completion.  No reference
 * counterpart. */
int pbx_debug_inject_wait_failure(int k);
/* Every tile update of a T-tile two-column factorisation (the launch plan
 * and the kernel's tile decoding, run on the host): records (launch, i, j,
 * m0, P), five ints each, of tile (i, j) less panels m0 .. m0+P-1, into rec
 * (up to cap records); returns the count.  own_cu: the one-workgroup-per-CU
 * build; K: the far period.  No reference counterpart: it lets a CPU test
 * check that every tile receives each panel once, in order, in time. */
int pbx_debug_two_col_updates(int T, int own_cu, int K, int* rec, int cap);
/* Kernels this library has launched in this process so far (every device,
 * every simulation).  Launches per step bound the step rate of small
 * simulations; no reference counterpart. */
long long pbx_kernel_launches(void);

/* C = A B^T (M x K times (N x K)^T); M, N, K multiples of 64.  Exposes the
 * FP64 MFMA tile product used by the factorisations, for testing. */
int pbx_gemm_nt(const double* A, const double* B, double* C, int M, int N, int K);

/* ------------------------------------------------------------------------
 * 3. Device-resident simulation: Box(beads, input_data) and
 *    Box.propagate(dt, build_Dff, build_Dnf, cholesky)
 *    (pyBrown/box.py:42, :150).  One HIP stream per handle; positions,
 *    matrices and factors stay in HBM between steps.
 * ------------------------------------------------------------------------ */

enum { PBX_NOHI = 0, PBX_RPY = 1, PBX_RPY_SMITH = 2, PBX_RPY_LUB = 3, PBX_RPY_SMITH_LUB = 4 };
enum { PBX_ERMAK = 0, PBX_MIDPOINT = 1 };

/* The JSON keywords of BD.py:37-64 that reach the step. */
typedef struct pbx_config {
    int hydrodynamics;          /* PBX_NOHI .. PBX_RPY_SMITH_LUB            */
    int scheme;                 /* propagation_scheme: PBX_ERMAK/MIDPOINT    */
    int ewald_real;             /* m_max                                     */
    int ewald_imag;             /* n_max                                     */
    int cichocki_correction;
    int check_overlaps;
    int measure_flux;
    int n_labels;               /* distinct bead labels (flux counters)     */
    long long max_move_attempts;
    int xi_pool_attempts;       /* attempts (3N draws) per noise chunk (0 = 64) */
    int factor_precision;       /* noise factor: 0 = FP64 (default), 1 = mixed:
                                   trailing updates on FP32 matrix cores,
                                   diagonal tiles and solves in FP64        */
    double box_length, T, viscosity;
    double ewald_alpha;
    double lubrication_cutoff;
    double m_midpoint;
    double overlap_treshold;    /* spelling of the reference keyword        */
    double flux_normal[3];
    double flux_plane_point[3];
    double external_force[3];
    int divergence_term;        /* divergence_term: the finite-difference
                                   divergence of D in lubrication Ermak steps
                                   (box.py:211-215, 560-610)                */
    int n_overlap_exempt;       /* pairs the overlap test skips (bonded      */
    const int* overlap_exempt;  /* beads: bead.py:401-413), 2 indices each   */
    int n_immobile;             /* immobile beads (box.py:649-668): fixed    */
    const double* immobile_xyz; /* positions (n_immobile x 3) and radii.     */
    const double* immobile_radii; /* They take bead indices N .. N+n_immobile-1
                                   in overlap_exempt and the force field, are
                                   tested for overlaps against mobile beads
                                   only, and are absent from D              */
} pbx_config;

typedef struct PbxSim pbx_sim;

/* Fills out[0..count) with successive N(0,1) draws; returns 0 on success.
 * The engine calls it, on the thread that calls pbx_sim_propagate, for one
 * chunk (xi_pool_attempts x 3N draws) at a time, ahead of use. */
typedef int (*pbx_xi_fill_fn)(void* user, double* out, long long count);

/* Box.__init__ (box.py:42-109): uploads radii, positions (N x 3) and label
 * indices; fails with PBX_ERR_OVERLAP (-5) for an overlapping start. */
int pbx_sim_create(const pbx_config* cfg, int N, const double* radii, const double* xyz, const int* label_idx,
                   pbx_sim** out);
int pbx_sim_destroy(pbx_sim* sim);
int pbx_sim_set_noise_source(pbx_sim* sim, pbx_xi_fill_fn fn, void* user);
/* Native noise source: the stream numpy.random.RandomState(seed).normal(0, 1)
 * produces (box.py:393-399, :621-629), generated ahead of use on a producer
 * thread.  Either call discards draws taken from a previous source. */
int pbx_sim_set_seed(pbx_sim* sim, unsigned int seed);
/* Move attempts evaluated per batch in the Ermak scheme with overlap checks
 * (1..32; 0 = adaptive, the default).  Trajectories do not depend on it. */
int pbx_sim_set_attempt_batch(pbx_sim* sim, int K);
/* Launch sequences with fixed arguments -- the noise factorisation and the
 * SPD inverses of the lubrication modes -- captured into HIP graphs on first
 * use and replayed as one graph launch per sequence afterwards (on = 1), or
 * launched kernel by kernel (on = 0, the default).  Results are identical.
 * No reference counterpart: it cuts host launch overhead where many small
 * simulations share the GPU (the serving loop). */
int pbx_sim_set_graphs(pbx_sim* sim, int on);
/* Lubrication modes: precision of the two SPD inverses behind each D
 * (box.py:514,524,537): 0 = FP64 (default), 1 = FP32 matrix-core bulk as in
 * pbx_spd_inverse_mixed, 2 = that plus one Newton-Schulz step (one more
 * 3N x 3N matrix of device memory).  No reference counterpart. */
int pbx_sim_set_inverse_precision(pbx_sim* sim, int precision);

/* Look-ahead assembly (on by default).  In the RPY and Smith-Ewald modes
 * (not with lubrication, not with HIP graphs on), a step that built D ends
 * by assembling D for its final positions into a spare matrix, after its
 * result copies are queued and while the host reads them; the host waits
 * for the results only.  The next step, when it builds D and the positions
 * were not replaced (pbx_sim_set_positions, pbx_sim_restore) in between,
 * takes that matrix instead of assembling.  Results are the same to the bit
 * either way; pbx_sim_get_matrix(D) always returns the D the last step used.
 * No counterpart in the reference (box.py:186-198 assembles at the start of
 * each step); for callers that step one simulation at a time. */
int pbx_sim_set_lookahead(pbx_sim* sim, int on);
/* Checkpoint: the device state (positions, flux, D, B[, Rff], diagonal
 * inverses, counters) and the noise state (unconsumed draws; the native
 * generator's state) as one byte buffer; restoring it into a simulation
 * created with the same configuration resumes bit for bit (pickling a Box,
 * test_box.py:53-117).  With a noise callback the caller restores its own
 * generator and registers the callback before pbx_sim_restore. */
long long pbx_sim_snapshot_size(pbx_sim* sim);
int pbx_sim_snapshot(pbx_sim* sim, void* buf, long long bytes);
int pbx_sim_restore(pbx_sim* sim, const void* buf, long long bytes);
/* count values of RandomState(seed).normal(0, 1) after skipping `skip`. */
int pbx_rng_normals(unsigned int seed, long long skip, double* out, long long count);

/* The step's force field (box.py:404-461, 685-723; interactions.py:483-589),
 * recomputed from the positions at the start of every propagate and used in
 * the deterministic terms D F dt / kBT (and the midpoint's (D_mid - D0) F):
 *   F = unit * (external_force, zeroed on beads outside an enabled axis range
 *               + Lennard-Jones over all bead pairs + harmonic bonds + angles),
 *   E = sum of the interaction energies (input energy unit; pbx_sim_get_energy).
 * Bonds and angles list bead indices (0-based, the order of the reference's
 * loops); angles in degrees.  Set once, after pbx_sim_create.  With immobile
 * beads (pbx_config.n_immobile) indices N.. name them: they act on mobile
 * beads (Lennard-Jones partners, bond and angle members, with the pair
 * vectors of interactions.py:185-416) and receive no force. */
typedef struct pbx_forces {
    int lennard_jones_6, lennard_jones_12;
    double lennard_jones_alpha;
    const double* hard_core_radius;  /* N + n_immobile (sigma = sum of the pair's), or NULL */
    const double* epsilon_LJ;        /* N + n_immobile (epsilon = sqrt of the product), or NULL */
    int n_bonds;
    const int* bond_beads;           /* 2 per bond: i (the bead listing it), j */
    const double* bond_params;       /* 2 per bond: equilibrium distance, force constant */
    int n_angles;
    const int* angle_beads;          /* 3 per angle: i, j, k */
    const double* angle_params;      /* 2 per angle: equilibrium angle (deg), force constant */
    int region_on[3];                /* external_force_region: axis x, y, z enabled */
    double region[6];                /* per axis: lower, upper bound (inclusive) */
    double unit;                     /* energy_unit factor on F: 1 (J), 6.9477e-21 (kcal/mol),
                                        1.602176634e-19 (eV) */
} pbx_forces;
int pbx_sim_set_forces(pbx_sim* sim, const pbx_forces* forces);
int pbx_sim_get_energy(pbx_sim* sim, double* E);
/* The force vector F (3N) the last step used (Box.F). */
int pbx_sim_get_force(pbx_sim* sim, double* F);

/* Box.propagate (box.py:150-198).  Returns 0; k > 0 for a non-SPD matrix
 * (pivot k-1); -4 when max_move_attempts is exhausted; other negatives for
 * argument/HIP errors. */
int pbx_sim_propagate(pbx_sim* sim, double dt, int build_Dff, int build_Dnf, int cholesky);

/* Box.propagate followed by the read-back the driver does after every step
 * (BD.py:168-175): positions (N x 3) and per-label flux counters, copied in
 * the step's own final synchronisation.  Either output may be NULL. */
int pbx_sim_propagate_out(pbx_sim* sim, double dt, int build_Dff, int build_Dnf, int cholesky, double* xyz_out,
                          long long* flux_out);

/* The BD.py:164-185 driver loop without file output, in one call: steps
 * j = j0 .. j0+steps-1 of Box.propagate with the cadence flags of BD.py:177
 * (j % diff_freq == 0, j % lub_freq == 0, j % chol_freq == 0).  flux_out
 * (steps x n_labels, or NULL) receives each step's per-label flux (the .flx
 * line after it); xyz_out (or NULL) the positions before each step j with
 * j % xyz_every == 0, one N x 3 frame per such step (the .xyz frames).
 * Results equal the same steps through pbx_sim_propagate_out. */
int pbx_sim_run(pbx_sim* sim, double dt, long long j0, long long steps, int diff_freq, int lub_freq, int chol_freq,
                long long* flux_out, double* xyz_out, int xyz_every);
/* Device memory held by the simulation, in bytes (what admission control
 * counts against an HBM budget). */
long long pbx_sim_device_bytes(pbx_sim* sim);
/* ||D - B B^T||_F / ||D||_F for the simulation's current diffusion matrix
 * and noise factor, computed on the device (FP64).  No reference
 * counterpart: the accuracy measure of the noise factor. */
int pbx_sim_factor_residual(pbx_sim* sim, double* rel);

int pbx_sim_get_positions(pbx_sim* sim, double* xyz);
int pbx_sim_set_positions(pbx_sim* sim, const double* xyz);
int pbx_sim_get_flux(pbx_sim* sim, long long* per_label);
/* which: 0 = D, 1 = B (lower, upper zeroed), 2 = far-field resistance Rff */
int pbx_sim_get_matrix(pbx_sim* sim, int which, double* out);
/* hydrodynamics nohi: which 0 = diagonal D, 1 = sqrt(D) (3N values) */
int pbx_sim_get_diag(pbx_sim* sim, int which, double* out);
long long pbx_sim_draw_count(pbx_sim* sim);
int pbx_sim_last_attempts(pbx_sim* sim);
int pbx_sim_sync(pbx_sim* sim);
/* Stage timing with HIP events on the simulation's stream (no extra sync
 * when off).  get_timing fills stage_ms[3] = cumulative milliseconds in
 * assembly (+lubrication), Cholesky, propagation; returns timed steps.
 * on = 1 records every stage bound (six events per step); on = 2 only the
 * two around the factorisation (stage_ms[0] and [2] stay zero). */
int pbx_sim_set_timing(pbx_sim* sim, int on);
long long pbx_sim_get_timing(pbx_sim* sim, double* stage_ms);

/* ------------------------------------------------------------------------
 * 4. Errors
 * ------------------------------------------------------------------------ */
const char* pbx_last_error(void);
int pbx_last_error_code(void);
void pbx_clear_error(void);
const char* pbx_version(void);

#ifdef __cplusplus
}
#endif
#endif
