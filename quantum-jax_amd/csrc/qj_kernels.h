// Internal host-side launcher interface between the C-ABI layer and the
// kernel translation units.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qj {

// Reals per cell of the acceleration grid A3 = (a_x, a_y, a_z): the
// reference's a_grid stacked on the last axis (quantum-stars.py:256-259).
constexpr int kAccComp = 3;

struct Dims {
  int nx, ny, nz;
  double Lx, Ly, Lz;
};

// Per-plan device tables the spectral passes read.
template <class T>
struct SpecTables {
  const void* tw_x;  // W_nx^p, p < nx   (cpx<T>)
  const void* tw_y;
  const void* tw_z;
  void* drift_x;     // exp(-i dt k^2/(2 mph)) / n per axis (cpx<T>)
  void* drift_y;
  void* drift_z;
  // schedule choices of the plan (results do not depend on them)
  int wide = 1;      // wide form of the Poisson x and y gradient passes
  int zblocks = 0;   // walking blocks of the z sweeps (0: the default cap)
};

// Poisson source: rho = |psi|^2 + src (either may be null); source_x also
// performs the x drift sweep on psi when drift != 0.
template <class T>
hipError_t launch_twiddles(void* tw, int n, hipStream_t s);
template <class T>
hipError_t launch_drift_table(void* tab, int n, double L, double c, hipStream_t s,
                              const double* dslot = nullptr);
// The three axes' tables (tab[a] of n[a] entries over box length L[a]) in one
// launch; dslot as in launch_drift_table.
template <class T>
hipError_t launch_drift_tables3(void* const tab[3], const int n[3], const double L[3], double c,
                                hipStream_t s, const double* dslot);

template <class T>
hipError_t launch_kick(void* psi, const void* V, const void* Vb, const void* Vx, double coef,
                       long long n, hipStream_t s);

// conj_d = 1 applies the conjugate phase table (the drift of -dt).
template <class T>
// psi_out (z axis only): the sweep reads psi and writes psi_out.
hipError_t launch_drift_axis(const Dims& d, const SpecTables<T>& tb, int axis, void* psi,
                             const void* V, const void* Vb, const void* Vx, double kick_coef,
                             int kick, hipStream_t s, int conj_d = 0,
                             const double* dslot = nullptr, void* psi_out = nullptr);

// The star deposit fused into the source pass: the column tiles (of
// make_column_tiles with the pass's workgroup columns, source_x_columns)
// hold the registrations of the stars (a scanned and scattered StarBins with
// the corner descriptors and weights); the pass adds their density to
// |psi|^2 and, with rho non-null, writes it to rho, every cell, as
// launch_tile_deposit would.
struct StarDeposit {
  const int* offset;
  const int* sorted;
  const unsigned long long* desc;  // StarBins::desc
  const void* wts;                 // StarBins::wts
  void* rho;                       // the density is written here, or nowhere (null)
  int ly, lz;  // the column tiles' TileGeom::ly and ::lz
  int ntiles;
  int home;    // StarBins::home
};

template <class T>
// psi_out (with drift): the drifted psi goes to psi_out instead of psi.
// dep: the fused star deposit (src must then be NULL).
hipError_t launch_source_x(const Dims& d, const SpecTables<T>& tb, int drift, void* psi,
                           void* src, int clear_src, void* brho, double b_coeff, double rho_bar,
                           void* H, double* partials, double* diag, hipStream_t s,
                           void* psi_out = nullptr, const StarDeposit* dep = nullptr);

// Columns per workgroup of the source pass on grid d in the precision of
// the c64 flag: the column-tile width of its fused deposit.
int source_x_columns(const Dims& d, int c64);

template <class T>
hipError_t launch_fft_y_half(const Dims& d, const SpecTables<T>& tb, void* H, hipStream_t s);

template <class T>
// partials: the per-block FP64 source sums of the preceding launch_source_x;
// the pass writes their compensated total to diag[kDiagMass].
hipError_t launch_poisson_z(const Dims& d, const SpecTables<T>& tb, int grad, void* H, void* Uz,
                            double four_pi_G, double rho_bar, int dc_mode,
                            const double* partials, double* diag, hipStream_t s,
                            double lp_kc2 = 0.0);

template <class T>
hipError_t launch_poisson_y(const Dims& d, const SpecTables<T>& tb, int grad, void* H, void* Uz,
                            void* Uy, hipStream_t s);

// The CFL signal folded into the Poisson x pass of a gas step: with the
// acceleration of each cell in registers the pass also loads the gas
// velocity there and writes, per workgroup, max over its cells of
// cs + |v + h a| (h = *owed, the kick owed after the step) to part[block];
// nblk receives the number of workgroups; a launch that would need more
// than cap entries of part is refused.
struct GasSignalFold {
  const void* v[3];
  const double* owed;
  double cs;
  double* part;
  int cap;
  int nblk;
};

// Fused star gather of a Poisson x gradient pass: workgroup t writes, for
// every star registered in column tile t (offset/sorted of the step's
// deposit over make_column_tiles(d, poisson_x_gather_columns(...))), the
// acceleration (a_x, a_y, a_z) of each of the star's CIC corners that lie in
// its columns into cv[(8 p + k) * 3 + c] (k: the corner order of the
// gather), locating them by the registrations' corner descriptors (desc, of
// the deposit's scatter pass).  A3 is then not written.
struct StarGather {
  const int* offset;
  const int* sorted;
  const unsigned long long* desc;  // corner descriptors of the registrations
  void* cv;
  int lz;    // log2 of the column tiles' z extent (their TileGeom::lz)
  int home;  // StarBins::home: lists of the lower y row, absolute descriptors
};

// Writes V and, with grad, the cell-interleaved grid A3 = (a_x, a_y, a_z)
// and folds max |a| into diag[kDiagAMax]; sig: the CFL signal fold above.
template <class T>
hipError_t launch_poisson_x(const Dims& d, const SpecTables<T>& tb, int grad, const void* H,
                            const void* Uy, const void* Uzz, void* V, void* A3,
                            const void* addend, double* diag, hipStream_t s,
                            GasSignalFold* sig = nullptr, const StarGather* gather = nullptr);

// Largest number of workgroups launch_poisson_x uses on grid d (either
// precision, either sweep form): the size of a GasSignalFold's part array.
int poisson_x_blocks(const Dims& d);

// Number of workgroups source_x launches (size of the FP64 partials array).
int source_x_blocks(const Dims& d);

// ---- particles (the grid acceleration is A3 = (a_x, a_y, a_z) per cell)
template <class T>
hipError_t launch_cic_indices(const Dims& d, const double* pos, long long np, int* i, int* ip1,
                              double* w_i, double* w_ip1, hipStream_t s);
template <class T>
hipError_t launch_deposit(const Dims& d, const double* pos, const double* mass, double m_scalar,
                          long long np, void* rho, hipStream_t s);
template <class T>
hipError_t launch_interp(const Dims& d, const double* pos, long long np, const void* A3,
                         double* acc, hipStream_t s);
// One star step up to the deposit: acc = interp(A3, pos);
// vel += acc*h_owed (the previous step's closing half kick); vel += acc*h1;
// pos = mod(pos + vel*dt, L); deposit(pos) into rho.
template <class T>
hipError_t launch_star_step(const Dims& d, double* pos, double* vel, long long np, const void* A3,
                            double h_owed, double h1, double dt, const double* mass,
                            double m_scalar, void* rho, hipStream_t s);
// vel += interp(A3)*h; diag[dt_s] = min |dx/vel| (diag[dt_s] must hold +inf)
template <class T>
hipError_t launch_star_kick(const Dims& d, const double* pos, double* vel, long long np,
                            const void* A3, double h, double* diag, hipStream_t s,
                            const void* cv = nullptr);

// ---- adjoints of the particle operations ----------------------------------
// Adjoint of vel += h*interp(A3, pos), preceded (dt_xv != 0) by that of
// pos += dt_xv*vel: vbar += dt_xv*xbar; then abar3[c].xyz += w_c*h*vbar
// (atomics; abar3 is (nx,ny,nz,3)) and xbar += grad terms.
template <class T>
hipError_t launch_interp_adj(const Dims& d, const double* pos, long long np, const void* A3,
                             double h, double* vbar, double dt_xv, double* xbar, void* abar3,
                             hipStream_t s);
// Adjoint of the deposit: mbar += interp(rhobar)/vol, xbar += grad terms.
template <class T>
hipError_t launch_deposit_adj(const Dims& d, const double* pos, const double* mass,
                              double m_scalar, long long np, const void* rhobar, double* mbar,
                              double* xbar, hipStream_t s);

// ---- adjoint spectral passes -------------------------------------------------
// psi_bar <- exp(+i coef V) psi_bar; Vbar = coef * Im(conj(psi_bar) psi_ref)
template <class T>
hipError_t launch_kick_adj(void* psi_bar, const void* psi_ref, const void* V, const void* Vx,
                           double coef, void* Vbar, long long n, hipStream_t s);
// psi_bar += 2 rhobar psi_ref
template <class T>
hipError_t launch_rho_adj(void* psi_bar, const void* psi_ref, const void* rhobar, long long n,
                          hipStream_t s);
// One axis of the drift adjoint (conjugate phase) on psi_bar, with
//   x or y axis and rhobar given: the density adjoint psi_bar += 2 rhobar
//     rho_ref applied to the rows before the transform (as launch_rho_adj);
//   z axis and Vbar given: the kick adjoint of the next half step applied
//     after it (as launch_kick_adj with V, Vx, coef and psi_ref = kick_ref).
// Without them it is the plain conjugate drift of launch_drift_axis.
template <class T>
hipError_t launch_drift_adj_axis(const Dims& d, const SpecTables<T>& tb, int axis, void* psi_bar,
                                 const void* rhobar, const void* rho_ref, const void* V,
                                 const void* Vx, double coef, const void* kick_ref, void* Vbar,
                                 hipStream_t s);
// rhobar = P(Vbar + div abar), P the Poisson operator (kernel -4 pi G/k^2 with
// the k = 0 bin -4 pi G for dc_mode 0, 0 for dc_mode 1), div the spectral
// divergence with the Nyquist derivative zeroed; abar3 is (nx,ny,nz,3).
template <class T>
hipError_t launch_adj_poisson(const Dims& d, const SpecTables<T>& tb, const void* Vbar,
                              const void* abar3, double four_pi_G, int dc_mode, void* H, void* U2,
                              void* U3, void* rhobar, hipStream_t s);

// ---- gas ----------------------------------------------------------------
// Device step slot of the CFL-limited stepper (doubles): the reduced signal
// speed, this step's dt, its owed + opening kick time, and the kick owed
// after it.  Kernels given the slot read dt and the kick time from it.
constexpr int kSlotCmax = 0, kSlotDt = 1, kSlotKick = 2, kSlotOwed = 3, kSlotCount = 4;
constexpr int kCflChunk = 64;  // CFL steps per host read-back

// Applies v += a*kick first when kick != 0 (A3 may be null when kick == 0).
// flux_kind: QJ_FLUX_RUSANOV (the reference's) or QJ_FLUX_HLL.  dslot
// (optional): dt and kick come from the device step slot instead.
template <class T>
hipError_t launch_hydro(const Dims& d, const void* rho, const void* vx, const void* vy,
                        const void* vz, const void* A3, double kick, double dt, double cs,
                        int flux_kind, void* rho_o, void* vx_o, void* vy_o, void* vz_o,
                        hipStream_t s, const double* dslot = nullptr);
template <class T>
hipError_t launch_gas_kick(long long n, void* vx, void* vy, void* vz, const void* A3, double h,
                           hipStream_t s);
// out[kSlotCmax] = max over cells of cs + |v + h a| (a from A3, skipped
// when A3 is null), out[kSlotDt] = min(dt_max, cfl dx / (Cmax + 1e-30)).
// With dts, out is a whole step slot: h is read from out[kSlotOwed], the
// slot advances (kick = owed + dt/2, owed = dt/2) and dts[k] = dt.  part:
// scratch for max_blocks per-block maxima.
template <class T>
hipError_t launch_gas_cfl(long long n, const void* vx, const void* vy, const void* vz,
                          const void* A3, double h, double cs, double dx, double cfl,
                          double dt_max, double* out, double* dts, int k, double* part,
                          int max_blocks, hipStream_t s);
// The reduction of launch_gas_cfl over signal maxima already in part[nblk]
// (a GasSignalFold), advancing the step slot (dts[k] = dt).
hipError_t launch_cfl_finish(const double* part, int nblk, double cfl, double dx, double dt_max,
                             double* slot, double* dts, int k, hipStream_t s);

// Tile-binned star deposit (owner computes; no global float atomics).
constexpr int kTileX = 8, kTileY = 8, kTileZ = 64;
constexpr int kTileMaxCells = kTileX * kTileY * kTileZ;
struct TileGeom {
  int lx, ly, lz;     // log2 of the tile extents
  int ntx, nty, ntz;  // tiles per axis
  int ntiles;
};
// Tile t's registration counter is count[t * kCountStride]: one 128-byte
// line per counter.  Counters that share a line share its atomic unit, and
// the counting atomics of a step (one to eight per star) then queue behind
// each other (32 counters per line made the counting pass of 2^16 stars on
// 512 tiles take 46 us instead of 16).
constexpr int kCountStride = 32;
struct StarBins {
  int* count;   // ntiles counters, kCountStride ints apart, zero between steps
  int* offset;  // ntiles + 1
  int* sorted;  // 8 np registrations at most
  int* rank;    // 8 np: slot of each registration within its tile
  int ntiles;
  TileGeom tg;  // the tiles the registrations refer to
  void* cv;     // corner values of the fused gather, 24 reals per star (or null)
  unsigned long long* desc;  // column tiles: corner descriptor per registration
  // column tiles: each star's eight corner weights m w / vol (8 reals per
  // star, corner (sx, sy, sz) at 4 sy + 2 sz + sx), written by the scatter
  // for the deposit fused into the source pass
  void* wts;
  // column tiles, set by the caller: 1 = each star registers only in the
  // tiles of its lower y row (one per z tile of its z pair), with absolute
  // corner descriptors (home_desc); a tile's consumers read its own list and
  // the list of the tile row below (the fused gather and deposit).  0 =
  // every tile its stencil touches, descriptors relative to the tile.
  int home;
};
TileGeom make_tiles(const Dims& d);
// Column tiles of the fused star gather: a tile is the set of x pencils that
// one workgroup of the Poisson x gradient pass transforms, all of x by cols
// consecutive (y, z) columns (q = y nz + z), so tile t is that pass's
// workgroup t and a step's deposit registrations are also its gather lists.
constexpr int kTileColMaxCells = 8192;
// rows: the tile's cols columns as rows y rows of cols / rows z columns
// (at most nz; rows 2 halves the registrations of a star's y pair).
TileGeom make_column_tiles(const Dims& d, int cols, int rows);
// Columns per workgroup of the Poisson x gradient pass on grid d (c64: the
// single-precision build; wide: the wide sweep form), or 0 when the fused
// gather does not serve that pass (column tiles above kTileColMaxCells).
int poisson_x_gather_columns(const Dims& d, int c64, int wide);
// Bytes of the registration part of a star scratch (either tile kind), and
// of the whole scratch with the corner-value area for reals of real_bytes.
size_t star_scratch_bytes(const Dims& d, long long np);
size_t star_scratch_total_bytes(const Dims& d, long long np, int real_bytes);
// cols = 0: the deposit's own tiles; else the column tiles of cols columns.
// real_bytes = 0: no corner-value area (a taped copy of the registrations).
StarBins carve_star_scratch(const Dims& d, long long np, void* scratch, int real_bytes = 0,
                            int cols = 0, int rows = 1);
// Kick/drift/wrap as launch_star_step (move = 1) or nothing (move = 0), and
// registration counting of the (new) positions in the tiles of sb.  The
// kicks interpolate A3, or with cv non-null the star's eight corner values
// that the fused gather left there (the same numbers, so the same result).
template <class T>
hipError_t launch_star_step_binned(const Dims& d, double* pos, double* vel, long long np,
                                   const void* A3, double h_owed, double h1, double dt, int move,
                                   const StarBins& sb, hipStream_t s, const void* cv = nullptr);
// Scan, scatter and per-tile deposit: overwrites every cell of rho.  det:
// order-independent fixed-point accumulation (bitwise reproducible).
// rho NULL: the scan and scatter only, the scatter also writing the corner
// weights sb.wts (the deposit is then fused into the source pass,
// StarDeposit).
template <class T>
hipError_t launch_tile_deposit(const Dims& d, const double* pos, const double* mass,
                               double m_scalar, long long np, const StarBins& sb, void* rho,
                               hipStream_t s, int det = 0);

// Interpolation adjoint whose h vbar corner values go to abar3 through the
// tile-binned deposit (whole tiles written, no clearing needed).  taped: sb
// already holds the registrations of pos (offsets and sorted), so they are
// not counted again.
template <class T>
hipError_t launch_interp_adj_binned(const Dims& d, const double* pos, long long np,
                                    const void* A3, double h, double* vbar, double dt_xv,
                                    double* xbar, void* abar3, const StarBins& sb, bool taped,
                                    hipStream_t s);

// Accretion onto black hole 0 (quantum-bh.py:306-339), in place on psi and m_bh.
// Ambient Bondi sink (quantum-fluid-accretion.py:440-510) on the gas
// density, in place; bh = (M_eff, x, y, z, vx, vy, vz) in device memory;
// out: 8 doubles of device memory; fmax < 0 = no fractional cap.
template <class T>
hipError_t launch_bondi_sink(const Dims& d, void* rho, const void* vx, const void* vy,
                             const void* vz, const double* bh, double cs, double G, double dt,
                             double r_acc_mult, double lam, double fmax, double ann_lo,
                             double ann_hi, double* out, hipStream_t s);

template <class T>
hipError_t launch_bh_accretion(const Dims& d, void* psi, const double* pos, const double* vel,
                               double* m_bh, long long n_bh, double G, double c, double sigma,
                               double mph, double dt, hipStream_t s);

// Sets the message returned by qj_last_error() and returns code (qj_api.hip).
int set_error(int code, const char* msg);

}  // namespace qj
