/* C ABI of the gfx950 KDK Schroedinger-Poisson library (libqjamd.so).
 *
 * The reference (pmocz/quantum-jax) has no FFI: its operator surface is a set
 * of module-level Python functions over jnp arrays that read module globals.
 * Each entry point below replaces one of them (file:line cited); the globals
 * become the plan (grid, box, precision) and explicit scalar arguments.
 *
 * Conventions
 *  - Arrays are device pointers in C order (nx, ny, nz), z contiguous.
 *    "complex" = interleaved (re, im) of the plan's real type; "real" = the
 *    plan's real type (float for QJ_C64, double for QJ_C128).  Particle
 *    positions, velocities, masses and accelerations are always double,
 *    shape (np, 3) row-major.
 *  - The caller owns every buffer; the plan owns only twiddle/phase tables
 *    and a few bytes of reduction scratch.  Operators that need temporary
 *    fields take a caller-provided workspace of qj_workspace_size() bytes.
 *  - Every call is asynchronous on `stream` (a hipStream_t, NULL = default
 *    stream) and performs no host synchronisation.  Scalars the reference
 *    returns to the host (a_max, dt_s) are written to a caller-owned device
 *    array of QJ_DIAG_COUNT doubles (`diag`).
 *  - Return value 0 on success, negative on error (QJ_E*); the message of the
 *    last error on this host thread is returned by qj_last_error().  Nothing
 *    aborts.  A plan must not be used concurrently from two streams.
 *  - Grid sizes: each of nx, ny, nz a power of two in [8, 512].
 */
#ifndef QJ_API_H
#define QJ_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QJ_C64 0
#define QJ_C128 1

#define QJ_OK 0
#define QJ_EARG (-1)
#define QJ_EUNSUPPORTED (-2)
#define QJ_EHIP (-3)

/* slots of the device diagnostics array (doubles) */
#define QJ_DIAG_AMAX 0 /* running max |a_grid|        quantum-stars.py:260,327 */
#define QJ_DIAG_DTS 1  /* min |dx / vel|              quantum-stars.py:328     */
#define QJ_DIAG_MASS 2 /* FP64 sum of the last Poisson source over cells     */
#define QJ_DIAG_NAN 3  /* 1.0 once a non-finite source value was seen        */
#define QJ_DIAG_COUNT 8

typedef struct qj_plan qj_plan;

/* Mesh, spectral setup and precision: replaces the module globals of
 * quantum-stars.py:72-75,131-149 (nx,ny,nz, Lx,Ly,Lz, kx,ky,kz,k_sq) and the
 * x64 switch of :69-70.  k is generated inside the kernels (FFT order,
 * Nyquist bin negative as in :142-148). */
int qj_plan_create(qj_plan** plan, int nx, int ny, int nz, double Lx, double Ly, double Lz,
                   int dtype);
int qj_plan_destroy(qj_plan* plan);
size_t qj_workspace_size(const qj_plan* plan);
const char* qj_last_error(void);
int qj_version(void);

/* get_potential(rho) -> V; quantum-stars.py:189-193, quantum-fluid.py:171-175.
 * dc_mode 0: the k = 0 bin follows the reference (divided by 1, not zeroed),
 * evaluated from an FP64 sum of rho; dc_mode 1: zero-mean V. */
int qj_get_potential(qj_plan* plan, const void* rho, double rho_bar, double four_pi_G,
                     double b_coeff, int dc_mode, void* V, void* work, void* stream);

/* a_d = -Re IFFT(1j k_d V_hat) on the grid and V; quantum-stars.py:254-260,
 * quantum-fluid.py:268-272.  a_grid is (nx, ny, nz, 3) reals holding
 * (a_x, a_y, a_z) per cell (the reference's a_grid, stacked on the last
 * axis); V (nullable) is the (nx, ny, nz) potential.  max |a| is folded
 * into diag[QJ_DIAG_AMAX] (diag nullable). */
int qj_acceleration_grid(qj_plan* plan, const void* rho, double rho_bar, double four_pi_G,
                         int dc_mode, void* V, void* a_grid, double* diag, void* work,
                         void* stream);

/* The low-pass-filtered gravity of the black-hole force of
 * quantum-fluid-accretion.py:140-146,182-191,719-725: V_hat of
 * get_potential_bg(rho, rho_bg) (the k = 0 bin divided by 1) times
 * LP = exp(-(k^2 / (k_cut^2 + 1e-30))^4), then a = -Re IFFT(1j k V_hat LP)
 * (accel_from_Vhat) into a_grid ((nx,ny,nz,3) reals) and the filtered
 * potential into V (nullable).  Sample a_grid at the hole with qj_interp. */
int qj_acceleration_grid_lp(qj_plan* plan, const void* rho, double rho_bg, double four_pi_G,
                            double k_cut, void* V, void* a_grid, void* work, void* stream);

/* bh_bondi_step_ambient of quantum-fluid-accretion.py:440-510 on the gas
 * density rho (real N, in place; cubic grid): sink radius r_acc =
 * r_acc_mult dx, ambient rho_inf (and, with vx/vy/vz given, the gas velocity
 * for the Bondi-Hoyle-Lyttleton rate) averaged over the shell
 * [max(ann_lo r_acc, r_acc + dx/4), max(ann_hi r_acc, r_lo + dx/2)], target
 * dM/dt = 4 pi lam (G M)^2 rho_inf / (cs^2 + v_rel^2)^1.5, removal
 * dM_cap = min(dM/dt dt, fmax M_avail) (fmax < 0: the reference's None, no
 * fraction) spread with the weights exp(-R^2/(2 rK^2)) rho inside r_acc.
 * bh: 7 device doubles (M_eff, x, y, z, vx, vy, vz); out: 8 device doubles
 * (dMdt_target, rho_inf, r_B, dM_cap, r_acc, M_avail, dM_uncl, shell
 * cells).  Cells farther than the shell from the hole are not touched (the
 * reference clips every cell at 0, which leaves non-negative densities as
 * they are). */
int qj_bondi_sink(qj_plan* plan, void* rho, const void* vx, const void* vy, const void* vz,
                  const double* bh, double cs, double G, double dt, double r_acc_mult,
                  double lam, double fmax, double ann_lo, double ann_hi, double* out,
                  void* stream);

/* psi <- exp(-1j coef V) psi; coef = m_per_hbar*dt/2. quantum-stars.py:301,318 */
int qj_kick(qj_plan* plan, void* psi, const void* V, double coef, void* stream);

/* psi <- IFFT(exp(-1j dt k^2/(2 m_per_hbar)) FFT psi); quantum-stars.py:307-309 */
int qj_drift(qj_plan* plan, void* psi, double dt, double m_per_hbar, void* stream);

/* get_cic_indices_and_weights(pos); quantum-stars.py:200-209 (FP64, bit-exact
 * indices; i/ip1 int32 (np,3), weights double (np,3)). */
int qj_cic_indices(qj_plan* plan, const double* pos, long long np, int* i, int* ip1, double* w_i,
                   double* w_ip1, void* stream);

/* bin_particles(pos, m) / bin_stars(pos); quantum-bh.py:223-258,
 * quantum-stars.py:212-247.  mass may be NULL (then every particle has
 * m_scalar).  accumulate = 0 clears rho first. */
int qj_bin_particles(qj_plan* plan, const double* pos, const double* mass, double m_scalar,
                     long long np, void* rho, int accumulate, void* stream);

/* 8-corner CIC interpolation of a_grid ((nx, ny, nz, 3) reals) to the
 * particles; quantum-stars.py:262-287.  acc is double (np, 3). */
int qj_interp(qj_plan* plan, const double* pos, long long np, const void* a_grid, double* acc,
              void* stream);

/* get_acceleration(pos, rho) -> (acc, a_max); quantum-stars.py:250-288.
 * a_grid ((nx,ny,nz,3) reals) receives the grid acceleration. */
int qj_get_acceleration(qj_plan* plan, const double* pos, long long np, const void* rho,
                        double rho_bar, double four_pi_G, int dc_mode, double* acc, void* a_grid,
                        double* diag, void* work, void* stream);

/* apply_grav_accel(rho, vx, vy, vz, dt); quantum-fluid.py:267-278 (in place
 * on vx, vy, vz; a_grid ((nx,ny,nz,3) reals of scratch) receives the grid
 * acceleration). */
int qj_apply_grav_accel(qj_plan* plan, const void* rho, void* vx, void* vy, void* vz, double dt,
                        double rho_bar, double four_pi_G, int dc_mode, void* a_grid, double* diag,
                        void* work, void* stream);

/* Gas face fluxes.  QJ_FLUX_RUSANOV is the reference's (the flux of the
 * averaged state plus Rusanov dissipation, quantum-fluid.py:238-264);
 * QJ_FLUX_HLL is the HLL flux with Davis signal speeds, which the reference
 * does not have. */
#define QJ_FLUX_RUSANOV 0
#define QJ_FLUX_HLL 1

/* solve_hydro(rho, vx, vy, vz, dt); quantum-fluid.py:281-349 (cubic cells;
 * out-of-place: the *_o buffers must not alias the inputs); flux is one of
 * QJ_FLUX_*. */
int qj_solve_hydro(qj_plan* plan, const void* rho, const void* vx, const void* vy, const void* vz,
                   double dt, double cs, int flux, void* rho_o, void* vx_o, void* vy_o,
                   void* vz_o, void* stream);

/* ---- fused stepping: compute_step + update + fori_loop ----------------
 * quantum-stars.py:295-351,454 (stars), quantum-fluid.py:414-459 (gas) and
 * the particle-free step.  The step kind follows the state: stars when
 * np > 0, gas when gas[0] != NULL, plain otherwise.
 *
 * The second half-kick of a step and the first of the next one use the same
 * potential (|psi|^2, star positions and gas density do not change between
 * them), so the state is kept with pending kicks: after qj_kdk_steps the
 * psi, gas-velocity and star-velocity buffers lack the last closing half
 * kick, which qj_kdk_finalize applies (and then sets diag[QJ_DIAG_DTS]).
 */
typedef struct {
  void* psi;        /* complex N                                              */
  void* rho_s;      /* real N, star density scratch                          */
  void* gas[4];     /* rho, vx, vy, vz: real N each (gas step) or NULL        */
  void* gas_tmp[4]; /* second buffers for the out-of-place gas update         */
  double* pos;      /* (np,3)                                                 */
  double* vel;      /* (np,3)                                                 */
  const double* mass; /* (np) or NULL                                         */
  double m_scalar;  /* star mass when mass == NULL                            */
  long long np;
  void* grid;       /* V (N reals) then A3 (N x 3 reals: a_x, a_y, a_z)        */
  void* brho;       /* real N, b*rho for self-interaction; NULL when b == 0   */
  double* diag;     /* QJ_DIAG_COUNT doubles                                  */
  /* host-side bookkeeping, initialise to 0 */
  int grid_valid;   /* 1: grid holds V and a of the current state; 2: V, and
                       the stars' corner accelerations are in star_scratch
                       (QJ_OPT_STAR_GATHER); 0: not computed                   */
  int gas_swapped;  /* the live gas fields are gas_tmp (after an odd count)   */
  double pending;   /* time of the owed kick of psi and gas velocity          */
  double pending_star; /* time of the owed closing kick of the star velocities */
  /* optional: qj_star_scratch_size() bytes.  When given (and
   * QJ_OPT_BINNED_DEPOSIT is on), the star deposit is tile-binned: per-tile
   * LDS accumulation and plain stores, no global float atomics; otherwise
   * the deposit adds with float atomics into rho_s after clearing it. */
  void* star_scratch;
  /* optional: real N, a static external potential added to V in the psi kick
   * (V = get_potential(rho) + external_potential(), quantum-tidal.py:175-202);
   * psi-only states */
  void* v_ext;
  /* host-side bookkeeping, initialise to 0: 1 when rho_s holds the star
   * density of pos; stars steps with the fused deposit that does not write
   * it (QJ_OPT_FUSED_DEPOSIT 2) leave 0, and qj_star_density computes it */
  int rho_valid;
} qj_state;

typedef struct {
  double dt;
  double m_per_hbar;
  double four_pi_G;
  double rho_bar;
  double b_coeff;
  double cs;
  int dc_mode;
  /* gas face flux: QJ_FLUX_RUSANOV (quantum-fluid.py:238-264) or QJ_FLUX_HLL */
  int flux;
} qj_params;

int qj_kdk_steps(qj_plan* plan, qj_state* state, const qj_params* prm, int nsteps, void* work,
                 void* stream);
/* qj_kdk_steps without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The workspace is a static device array of
 * QJ_NOALLOC_SCRATCH_BYTES per GPU, looked up when the first plan on that
 * GPU is created and held until the process ends; the call itself allocates
 * nothing.  Plans whose qj_workspace_size() exceeds it (e.g. 256^3 c64;
 * 128^3 c64 fits) are rejected with QJ_EUNSUPPORTED, as is a plan with stage
 * timing on.  Every call on one GPU shares that array: calls that overlap in
 * time on one GPU (two streams, two threads) race and must not be issued. */
#define QJ_NOALLOC_SCRATCH_BYTES (40u * 1024u * 1024u)
int qj_kdk_steps_noalloc(qj_plan* plan, qj_state* state, const qj_params* prm, int nsteps,
                         void* stream);
/* nsteps steps with per-step kick and drift times (host arrays of nsteps):
 * the kicks of step k use kick_dt[k]/2 and the drift (and the gas update)
 * drift_dt[k].  The comoving step of quantum-cosmo.py:166-197 is
 * kick_dt = dt pot_fac and drift_dt = dt kin_fac; a CFL-limited gas run
 * (qj_gas_cfl_dt) passes kick_dt = drift_dt = its dt.  Psi and gas states
 * (no stars). */
int qj_kdk_steps_sched(qj_plan* plan, qj_state* state, const qj_params* prm, int nsteps,
                       const double* kick_dt, const double* drift_dt, void* work, void* stream);

/* CFL time step of the gas state (quantum-fluid-accretion.py:696-701,
 * isothermal: c_eff = cs): out[0] = max over cells of cs + |v|, with v the
 * live velocity plus the owed half kick; out[1] = min(dt_max,
 * cfl dx / (out[0] + 1e-30)).  out: 2 doubles of device memory (a NaN
 * velocity gives NaN).  Stream-ordered; the caller reads out[1] to choose
 * the next step's dt. */
int qj_gas_cfl_dt(qj_plan* plan, const qj_state* state, const qj_params* prm, double cfl,
                  double dt_max, double* out, void* stream);

/* nsteps gas steps, each with dt = the qj_gas_cfl_dt value of the state it
 * starts from (quantum-fluid-accretion.py:696-701: the CFL cap is evaluated
 * inside every step).  dt is computed and consumed on the device (the drift
 * tables, the kicks and the gas update read it there); the host reads the
 * steps taken back once per 64 steps, so the call synchronises the stream.
 * dts (host, nsteps doubles, may be NULL) receives them.  A non-finite dt
 * (a non-finite velocity) returns an error after its chunk.  Psi and gas
 * states (no stars). */
int qj_kdk_steps_cfl(qj_plan* plan, qj_state* state, const qj_params* prm, int nsteps, double cfl,
                     double dt_max, double* dts, void* work, void* stream);
/* do_accretion(psi, pos, vel, m_bh) of quantum-bh.py:306-339, in place:
 * black hole 0 takes dM/dt dt from the field at its cell (the reference
 * lowers the amplitude there by sqrt(dM/dt dt / vol) at fixed phase) and
 * dM/dt dt is added to each of the n_bh masses (the reference adds the
 * scalar to its whole mass array).  pos, vel, m_bh: the black holes' FP64
 * arrays. */
int qj_bh_accretion(qj_plan* plan, void* psi, const double* pos, const double* vel, double* m_bh,
                    long long n_bh, double G, double c, double sigma, double m_per_hbar,
                    double dt, void* stream);

/* The star density rho_s of the state's current positions (the deposit of
 * quantum-stars.py:212-247, bin_stars), for a state whose rho_valid is 0:
 * the stars are registered again in the deposit's own tiles and added with
 * the deterministic (fixed-point, order-independent) tile deposit when the
 * state has a star scratch, else with float atomics.  Sets rho_valid.  The
 * registrations of the last step in star_scratch are overwritten (the next
 * step registers its own); the fused gather's corner values are kept. */
int qj_star_density(qj_plan* plan, qj_state* state, void* stream);

/* Bytes of qj_state.star_scratch for np stars: per-tile counters and
 * offsets and up to 8 registrations per star (bin_stars of
 * quantum-stars.py:212-247, restructured as owner-computes by tile). */
size_t qj_star_scratch_size(const qj_plan* plan, long long np);
int qj_kdk_finalize(qj_plan* plan, qj_state* state, const qj_params* prm, void* stream);

/* ---- adjoint of the KDK step (no counterpart in the reference, which
 * never differentiates despite README.md:7) ----------------------------
 * qj_kdk_record runs nsteps steps (quantum-stars.py:295-330; the state
 * must be finalized; the steps merge consecutive half kicks as qj_kdk_steps
 * does, and the state is finalized at the end) and stores, per step, psi
 * after the drift (psi2), the
 * grid after the step's Poisson solve (grid1, 4N reals as in qj_state), the
 * star positions after the drift (x1), and the star positions and psi at
 * the start (x0, psi0: needed for step 0 only, NULL allowed for later steps:
 * step k starts from x1 and psi2 of step k - 1); grid0
 * (nullable) receives the grid at the start.  The adjoint merges the
 * closing half kick of each step with the opening one of the next (one
 * adjoint Poisson solve per step).
 * qj_kdk_adjoint then maps the adjoint of the end state (psi_bar, x_bar,
 * v_bar; psi_bar = dL/dRe psi + i dL/dIm psi) back through those steps in
 * place and accumulates m_bar (dL/dm per star).  scratch holds
 * qj_adjoint_scratch_size() bytes.  Stars and plain steps only. */
typedef struct {
  void* psi0;
  void* psi2;
  void* grid1;
  double* x0;
  double* x1;
  /* optional (stars with the tile-binned deposit): qj_star_scratch_size()
   * bytes that qj_kdk_record fills with the tile registrations of the
   * positions x1 and qj_kdk_adjoint reuses instead of recounting them;
   * NULL = recount.  Keep QJ_OPT_BINNED_DEPOSIT unchanged between the two
   * calls: with it off the record leaves bins unwritten. */
  void* bins;
} qj_tape_step;

typedef struct {
  void* psi_bar;  /* complex N, in/out */
  double* x_bar;  /* (np,3), in/out   */
  double* v_bar;  /* (np,3), in/out   */
  double* m_bar;  /* (np), accumulated */
  void* scratch;
} qj_adjoint;

size_t qj_adjoint_scratch_size(const qj_plan* plan);
int qj_kdk_record(qj_plan* plan, qj_state* state, const qj_params* prm, int nsteps,
                  qj_tape_step* tape, void* grid0, void* work, void* stream);
int qj_kdk_adjoint(qj_plan* plan, const qj_state* state, const qj_params* prm, int nsteps,
                   const qj_tape_step* tape, const void* grid0, qj_adjoint* adj, void* work,
                   void* stream);
/* Adjoint of bin_particles: m_bar += interp(rho_bar)/vol, x_bar += the
 * position gradient (CIC weights are piecewise linear). */
int qj_deposit_adjoint(qj_plan* plan, const double* pos, const double* mass, double m_scalar,
                       long long np, const void* rho_bar, double* m_bar, double* x_bar,
                       void* stream);

/* ---- schedule options (no counterpart in the reference, whose XLA
 * compiler picks the fusion) -------------------------------------------
 * QJ_OPT_FUSE_X_SOURCE: 1 = the x drift sweep and the Poisson source pass
 * run as one kernel (fewer HBM bytes, more work per workgroup); 0 = two
 * kernels; 2 (default) = one kernel, except in stars steps whose star work
 * runs on the side stream with the fused deposit on x pencils of 256 points
 * or more, where the separate x drift runs beside the star work too.
 * Results are identical; the autotuner picks per grid. */
#define QJ_OPT_FUSE_X_SOURCE 0
/* QJ_OPT_BINNED_DEPOSIT: with qj_state.star_scratch given, 1 = tile-binned
 * star deposit (default), 0 = float-atomic deposit (cheaper for sparse
 * stars on small grids).  Results agree to rounding. */
#define QJ_OPT_BINNED_DEPOSIT 1
/* QJ_OPT_DETERMINISTIC_DEPOSIT: 1 = the tile-binned deposit adds the
 * corner values as 64-bit fixed-point integers per tile (order-independent
 * sums: stars runs reproduce bit for bit, and rho_s is within FP64 rounding
 * of the exact sum); 0 (default) = float adds in LDS, whose order varies
 * from run to run.  Needs the binned deposit (star_scratch given and
 * QJ_OPT_BINNED_DEPOSIT on). */
#define QJ_OPT_DETERMINISTIC_DEPOSIT 2
/* QJ_OPT_WIDE_SWEEPS: 1 (default) = the Poisson x and y gradient passes of
 * complex64 256-point pencils run one pencil per lane in 512-thread
 * workgroups; 0 = two pencils per lane in 256 threads.  Same results. */
#define QJ_OPT_WIDE_SWEEPS 3
/* QJ_OPT_Z_SWEEP_BLOCKS: workgroups of the z drift sweeps, which walk the
 * pencil groups (0 = default: at most 512, or at most 1024 when the side
 * stream runs star work beside the sweep on a grid of fewer than 256^3
 * cells; otherwise at most this many).
 * Same results. */
#define QJ_OPT_Z_SWEEP_BLOCKS 4
/* QJ_OPT_SIDE_STREAM (default 1): the stepper runs each step's star kick,
 * drift and deposit on a second stream of the plan, beside the psi drift
 * sweeps that do not depend on them, and joins the caller's stream before
 * the Poisson source pass; 0 = everything on the caller's stream.  Same
 * results. */
#define QJ_OPT_SIDE_STREAM 5
/* QJ_OPT_STAR_GATHER (default 1): in stars steps with the tile-binned
 * deposit (no gas, not recording a tape) the deposit bins the stars by the
 * columns of the Poisson x pass, and that pass hands each star the
 * accelerations at its eight CIC corners instead of writing the
 * acceleration grid; the next step's star kicks read them.  Same star and
 * psi results (the same corner values); the state's acceleration grid
 * (grid after V) is then not written (qj_state.grid_valid 2); 0 = the grid
 * is written and the stars gather from it. */
#define QJ_OPT_STAR_GATHER 6
/* QJ_OPT_GATHER_ROWS: y rows of the fused gather's column tiles (0 =
 * automatic, 1 or 2).  Two rows cut the stars' registrations by a quarter;
 * on grids where they leave the Poisson x pass half-line rows they pay only
 * with many stars per cell (automatic: at least one star per 32 cells, and
 * not with the fused deposit, whose stars register in one row anyway).
 * Same results to rounding. */
#define QJ_OPT_GATHER_ROWS 7
/* QJ_OPT_FUSED_DEPOSIT (default 2): where the fused gather's column tiles
 * serve a stars step and the deposit is not the deterministic one, the
 * Poisson source pass deposits the stars of its column tile itself (in LDS,
 * added to |psi|^2); 2 = rho_s is then not written (qj_state.rho_valid 0,
 * qj_star_density computes it when a caller needs it), 1 = the pass also
 * writes the tile's density to rho_s, 0 = the separate per-tile deposit
 * kernel writes rho_s and the source pass reads it.  With 1 and 2 each star
 * registers only in the column tile of its lower y row, and the source
 * pass and the gather also read the list of the tile row below.  Same
 * results to rounding (float atomics add in either order). */
#define QJ_OPT_FUSED_DEPOSIT 8
#define QJ_OPT_COUNT 9
int qj_plan_set_option(qj_plan* plan, int option, int value);
int qj_plan_get_option(const qj_plan* plan, int option, int* value);

/* ---- ensembles of independent boxes (no counterpart in the reference,
 * which runs one box per process: quantum-stars.py:454) -----------------
 * An ensemble steps many boxes (any mix of sizes, precisions, step kinds
 * and dt) together.  For each step it records the launches that
 * qj_kdk_steps would issue for every member, puts members whose launch
 * sequences have the same shape into one group and issues one batched
 * launch per stage and group (the group's members are blockIdx.z of the
 * launch).  The recorded step is cached and, in the graph modes, captured
 * once into a hipGraph that is replayed every step.  Results are bitwise
 * identical to qj_kdk_steps on each member.
 *
 * members: nmembers plans, states, params and workspaces (arrays).  Plans,
 * states and workspaces must be distinct per member, stay alive, and not be
 * used outside the ensemble while it lives.  A cached step is keyed by the
 * members' states, params and plan options, so changing any of them between
 * calls is allowed (it records anew). */
#define QJ_ENS_EAGER 0      /* member after member, ordinary launches       */
#define QJ_ENS_BATCHED 1    /* one batched launch per stage and group       */
#define QJ_ENS_GRAPH 2      /* batched launches replayed from a hipGraph    */
#define QJ_ENS_GRAPH_FORK 3 /* as 2, one concurrent graph branch per group  */
typedef struct qj_ensemble qj_ensemble;
int qj_ensemble_create(qj_ensemble** ens, int nmembers, qj_plan* const* plans,
                       qj_state* const* states, const qj_params* params, void* const* works);
int qj_ensemble_destroy(qj_ensemble* ens);
int qj_ensemble_set_mode(qj_ensemble* ens, int mode);
int qj_ensemble_steps(qj_ensemble* ens, int nsteps, void* stream);
int qj_ensemble_finalize(qj_ensemble* ens, void* stream);
/* qj_kdk_record and qj_kdk_adjoint of every member (tapes[m] holds nsteps
 * steps of member m; grid0s and its entries nullable; adjs[m] is member m's
 * adjoint), recorded, batched and replayed like qj_ensemble_steps. */
int qj_ensemble_record(qj_ensemble* ens, int nsteps, qj_tape_step* const* tapes,
                       void* const* grid0s, void* stream);
int qj_ensemble_adjoint(qj_ensemble* ens, int nsteps, qj_tape_step* const* tapes,
                        void* const* grid0s, qj_adjoint* adjs, void* stream);
/* Groups of the last recorded step: count, then per group its members
 * (written to members[0..cap)), member count and launches per step. */
int qj_ensemble_group_count(const qj_ensemble* ens);
int qj_ensemble_group_info(const qj_ensemble* ens, int group, int* members, int cap,
                           int* nmembers, int* launches);

/* ---- stage timing (tracing hook; no counterpart in the reference, which
 * only wall-clocks its chunk loop at quantum-stars.py:451,467) ----------
 * When enabled, every kernel launch of the plan is bracketed by a pair of
 * HIP events on its stream; qj_timing_read synchronises on them and returns
 * (then clears) the summed milliseconds and launch count per stage. */
int qj_stage_count(void);
const char* qj_stage_name(int stage);
int qj_timing_enable(qj_plan* plan, int enable);
int qj_timing_read(qj_plan* plan, double* ms, long long* launches, int nstages);

#ifdef __cplusplus
}
#endif

#endif /* QJ_API_H */
