// C-ABI layer: plan management, argument checking and the composition of
// the kernels into the reference operators and the fused KDK step.
#include <math.h>
#include <string.h>

#include <string>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/qj_api.h"
#include "qj_common.h"
#include "qj_kernels.h"

using namespace qj;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

int qj::set_error(int code, const char* msg) {
  g_err = msg;
  return code;
}

namespace {

int hipfail(hipError_t e, const char* where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return QJ_EHIP;
}

#define QJ_CHECK(expr, where)                    \
  do {                                           \
    hipError_t e_ = (expr);                      \
    if (e_ != hipSuccess) return hipfail(e_, where); \
  } while (0)

bool pow2_ok(int n) { return n >= 8 && n <= 512 && (n & (n - 1)) == 0; }

template <class T>
struct k_fill {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(T* a, long long n, T v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) a[i] = v;
  }
};

template <class T>
hipError_t fill(void* a, long long n, T v, hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  launch<k_fill<T>>((int)blocks, 256, s, (T*)a, n, v);
  return hipGetLastError();
}

// Device-to-device copy as a kernel (recordable by the launch recorder, so
// tape copies can be batched and graph-captured with the rest of a step).
template <class V>
struct k_copy {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(V* __restrict__ dst, const V* __restrict__ src,
                                             long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
  }
};

hipError_t copy_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const uintptr_t a = (uintptr_t)dst | (uintptr_t)src;
  if ((a | bytes) % 8 != 0) return hipErrorInvalidValue;  // every tape buffer is 8-byte data
  const bool wide = (a | bytes) % 16 == 0;
  const long long n = (long long)(bytes / (wide ? 16 : 8));
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (wide)
    launch<k_copy<uint4>>((int)blocks, 256, s, (uint4*)dst, (const uint4*)src, n);
  else
    launch<k_copy<uint2>>((int)blocks, 256, s, (uint2*)dst, (const uint2*)src, n);
  return hipGetLastError();
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Kernel stages that the optional event timer tells apart.
enum Stage {
  kStageDriftZKick = 0, kStageDriftZ, kStageDriftY, kStageDriftX, kStageDriftXSource,
  kStageSourceX, kStageFftY, kStagePoissonZ, kStagePoissonY, kStagePoissonX, kStageDeposit,
  kStageStarStep, kStageStarKick, kStageHydro, kStageGasKick, kStageKick, kStageAdjParticles,
  kStageAdjKick, kStageAdjPoisson, kStageAdjRho, kStageAdjDrift, kStageCount
};
const char* kStageNames[kStageCount] = {
    "drift_z_kick", "drift_z", "drift_y", "drift_x", "drift_x_source", "source_x", "fft_y_half",
    "poisson_z", "poisson_y", "poisson_x", "deposit", "star_step", "star_kick",
    "hydro", "gas_kick", "kick", "adj_particles", "adj_kick", "adj_poisson", "adj_rho",
    "adj_drift"};

}  // namespace

// Workspace of qj_kdk_steps_noalloc: a file-scope device array, one copy per
// GPU the library runs on, held until the process ends.  Host code never
// passes its name to a kernel or a copy (in host code the name is a host-side
// stand-in); it uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[QJ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device dev (which must be current), looked
// up once per device and kept.
int scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(QJ_EUNSUPPORTED, "no-allocation workspace: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hipfail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(QJ_EHIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return QJ_OK;
}

}  // namespace

struct qj_plan {
  Dims d;
  int dtype;
  int device;
  size_t rsize;  // bytes of one real
  void* tw[3] = {nullptr, nullptr, nullptr};
  void* drift[3] = {nullptr, nullptr, nullptr};
  double drift_key_dt = NAN, drift_key_mph = NAN;
  double* partials = nullptr;
  int nblk_src = 0;
  // per-workgroup CFL signal maxima of the folded Poisson x pass
  double* sig_part = nullptr;
  int nblk_sig = 0;
  // schedule options (qj_plan_set_option): x schedule (automatic), binned,
  // deterministic, wide Poisson sweeps, z-sweep blocks, side stream, star
  // gather, gather rows (automatic), fused deposit (without the rho_s write)
  int opt[QJ_OPT_COUNT] = {2, 1, 0, 1, 0, 1, 1, 0, 2};
  // second stream (default priority) for the star and gas work of a step,
  // with the fork and join events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // optional per-stage event timing (tracing hook of the bench harness)
  int timing = 0;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double stage_ms[kStageCount] = {};
  long long stage_n[kStageCount] = {};

  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }

  long long ncell() const { return (long long)d.nx * d.ny * d.nz; }
  long long nhalf() const { return (long long)(d.nx / 2 + 1) * d.ny * d.nz; }
};

namespace {

// Records a (begin, end) event pair around one launch when timing is on.
struct StageTimer {
  qj_plan* p;
  int id;
  hipStream_t s;
  hipEvent_t a = nullptr;
  StageTimer(qj_plan* p_, int id_, hipStream_t s_) : p(p_), id(id_), s(s_) {
    if (p->timing && !g_recorder) {
      a = p->take();
      if (a) (void)hipEventRecord(a, s);
    }
  }
  ~StageTimer() {
    if (!a) return;
    hipEvent_t b = p->take();
    if (!b) return;
    (void)hipEventRecord(b, s);
    p->recs.push_back({id, a, b});
  }
};

#define QJ_STAGE(id, expr, where)                      \
  do {                                                 \
    StageTimer st_(p, id, s);                          \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return hipfail(e_, where);   \
  } while (0)

struct Work {
  void* H;
  void* Uz;
  void* Uy;
  void* R;
};

Work carve(const qj_plan* p, void* work) {
  char* w = (char*)work;
  const size_t hb = align256(p->nhalf() * 2 * p->rsize);
  Work k;
  k.H = w;
  k.Uz = w + hb;
  k.Uy = w + 2 * hb;
  k.R = w + 3 * hb;
  return k;
}

template <class T>
SpecTables<T> tables(const qj_plan* p) {
  SpecTables<T> t;
  t.tw_x = p->tw[0];
  t.tw_y = p->tw[1];
  t.tw_z = p->tw[2];
  t.drift_x = p->drift[0];
  t.drift_y = p->drift[1];
  t.drift_z = p->drift[2];
  t.wide = p->opt[QJ_OPT_WIDE_SWEEPS];
  t.zblocks = p->opt[QJ_OPT_Z_SWEEP_BLOCKS];
  return t;
}

int check_plan(const qj_plan* p) {
  if (!p) return fail(QJ_EARG, "null plan");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != p->device)
    return fail(QJ_EARG, "current HIP device differs from the plan's device");
  return QJ_OK;
}

template <class T>
int ensure_drift(qj_plan* p, double dt, double mph, hipStream_t s) {
  if (p->drift_key_dt == dt && p->drift_key_mph == mph) return QJ_OK;
  const double c = dt / (2.0 * mph);  // quantum-stars.py:308: dt * k^2 / m_per_hbar / 2
  const int n[3] = {p->d.nx, p->d.ny, p->d.nz};
  const double L[3] = {p->d.Lx, p->d.Ly, p->d.Lz};
  for (int a = 0; a < 3; ++a)
    QJ_CHECK(launch_drift_table<T>(p->drift[a], n[a], L[a], c, s), "drift table");
  p->drift_key_dt = dt;
  p->drift_key_mph = mph;
  return QJ_OK;
}

// Poisson chain after the source pass: H -> grid (V [, a]).
template <class T>
int poisson_tail(qj_plan* p, const Work& w, int grad, double four_pi_G, double rho_bar,
                 int dc_mode, void* V, void* A3, const void* addend,
                 double* diag, hipStream_t s, double lp_kc2 = 0.0,
                 GasSignalFold* sig = nullptr, const StarGather* gather = nullptr) {
  SpecTables<T> tb = tables<T>(p);
  QJ_STAGE(kStageFftY, launch_fft_y_half<T>(p->d, tb, w.H, s), "fft_y");
  QJ_STAGE(kStagePoissonZ, launch_poisson_z<T>(p->d, tb, grad, w.H, w.Uz, four_pi_G, rho_bar, dc_mode,
                                                  p->partials, diag, s, lp_kc2),
           "poisson_z");
  QJ_STAGE(kStagePoissonY, launch_poisson_y<T>(p->d, tb, grad, w.H, w.Uz, w.Uy, s), "poisson_y");
  QJ_STAGE(kStagePoissonX,
           launch_poisson_x<T>(p->d, tb, grad, w.H, w.Uy, w.Uz, V, A3, addend, diag, s, sig,
                               gather),
           "poisson_x");
  return QJ_OK;
}

// Real plane k of a buffer of N-real planes (the grid is V then A3).
template <class T>
void* plane(void* base, const qj_plan* p, int k) {
  return (char*)base + (size_t)k * p->ncell() * sizeof(T);
}

// ---------------------------------------------------------------------------
template <class T>
int get_potential_t(qj_plan* p, const void* rho, double rho_bar, double four_pi_G, double b,
                    int dc_mode, void* V, void* work, hipStream_t s) {
  Work w = carve(p, work);
  SpecTables<T> tb = tables<T>(p);
  void* brho = (b != 0.0) ? w.R : nullptr;
  // diag scratch: the source mass lands in the plan-owned partials' tail
  double* diag = p->partials + p->nblk_src;
  QJ_STAGE(kStageSourceX, launch_source_x<T>(p->d, tb, 0, nullptr, (void*)rho, 0, brho, b, rho_bar, w.H,
                              p->partials, diag, s),
           "source_x");
  return poisson_tail<T>(p, w, 0, four_pi_G, rho_bar, dc_mode, V, nullptr, brho, diag, s);
}

template <class T>
int accel_grid_t(qj_plan* p, const void* rho, double rho_bar, double four_pi_G, int dc_mode,
                 void* V, void* a_grid, double* diag, void* work, hipStream_t s,
                 double lp_kc2 = 0.0) {
  Work w = carve(p, work);
  SpecTables<T> tb = tables<T>(p);
  // the source mass and max|a| land in the caller's diag, else in plan scratch
  double* dg = diag ? diag : p->partials + p->nblk_src;
  QJ_STAGE(kStageSourceX, launch_source_x<T>(p->d, tb, 0, nullptr, (void*)rho, 0, nullptr, 0.0, rho_bar, w.H,
                              p->partials, dg, s),
           "source_x");
  void* Vout = V ? V : w.R;  // the pass always writes V
  return poisson_tail<T>(p, w, 1, four_pi_G, rho_bar, dc_mode, Vout, a_grid, nullptr, dg, s,
                         lp_kc2);
}

}  // namespace

// ===========================================================================
// exported entry points
// ===========================================================================
extern "C" {

const char* qj_last_error(void) { return g_err.c_str(); }

int qj_version(void) { return 1; }

int qj_plan_create(qj_plan** out, int nx, int ny, int nz, double Lx, double Ly, double Lz,
                   int dtype) {
  if (!out) return fail(QJ_EARG, "null output pointer");
  *out = nullptr;
  if (!pow2_ok(nx) || !pow2_ok(ny) || !pow2_ok(nz))
    return fail(QJ_EUNSUPPORTED, "grid sizes must be powers of two in [8, 512]");
  if (!(Lx > 0) || !(Ly > 0) || !(Lz > 0)) return fail(QJ_EARG, "box lengths must be positive");
  if (dtype != QJ_C64 && dtype != QJ_C128) return fail(QJ_EARG, "dtype must be QJ_C64 or QJ_C128");
  qj_plan* p = new qj_plan();
  p->d = Dims{nx, ny, nz, Lx, Ly, Lz};
  p->dtype = dtype;
  p->rsize = dtype == QJ_C64 ? sizeof(float) : sizeof(double);
  hipError_t e = hipGetDevice(&p->device);
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    e = hipMalloc(&p->tw[a], (size_t)n[a] * 2 * p->rsize);
    if (e == hipSuccess) e = hipMalloc(&p->drift[a], (size_t)n[a] * 2 * p->rsize);
  }
  p->nblk_src = source_x_blocks(p->d);
  p->nblk_sig = poisson_x_blocks(p->d);
  const size_t nfixed = (size_t)p->nblk_src + kDiagCount + kSlotCount + kCflChunk;
  const size_t nscratch = nfixed + p->nblk_sig;
  if (e == hipSuccess) e = hipMalloc((void**)&p->partials, nscratch * sizeof(double));
  if (e == hipSuccess) p->sig_part = p->partials + nfixed;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMemset(p->partials, 0, nscratch * sizeof(double));
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    e = dtype == QJ_C64 ? launch_twiddles<float>(p->tw[a], n[a], nullptr)
                        : launch_twiddles<double>(p->tw[a], n[a], nullptr);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    qj_plan_destroy(p);
    return hipfail(e, "qj_plan_create");
  }
  // set-up of the no-allocation workspace on this device (once per device)
  void* sp = nullptr;
  if (int rc = scratch_address(p->device, &sp)) {
    qj_plan_destroy(p);
    return rc;
  }
  *out = p;
  return QJ_OK;
}

int qj_plan_destroy(qj_plan* p) {
  if (!p) return QJ_OK;
  for (auto& r : p->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : p->pool) (void)hipEventDestroy(e);
  for (int a = 0; a < 3; ++a) {
    if (p->tw[a]) (void)hipFree(p->tw[a]);
    if (p->drift[a]) (void)hipFree(p->drift[a]);
  }
  if (p->partials) (void)hipFree(p->partials);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->side) (void)hipStreamDestroy(p->side);
  delete p;
  return QJ_OK;
}

int qj_plan_set_option(qj_plan* p, int option, int value) {
  if (!p) return fail(QJ_EARG, "null plan");
  if (option < 0 || option >= QJ_OPT_COUNT) return fail(QJ_EARG, "unknown option");
  if (option == QJ_OPT_Z_SWEEP_BLOCKS && (value < 0 || value > 65535))
    return fail(QJ_EARG, "QJ_OPT_Z_SWEEP_BLOCKS must be in [0, 65535]");
  const bool three = option == QJ_OPT_GATHER_ROWS || option == QJ_OPT_FUSED_DEPOSIT ||
                     option == QJ_OPT_FUSE_X_SOURCE;
  if (three && (value < 0 || value > 2)) return fail(QJ_EARG, "this option takes 0, 1 or 2");
  if (option != QJ_OPT_Z_SWEEP_BLOCKS && !three && value != 0 && value != 1)
    return fail(QJ_EARG, "this option takes 0 or 1");
  p->opt[option] = value;
  return QJ_OK;
}

int qj_plan_get_option(const qj_plan* p, int option, int* value) {
  if (!p || !value) return fail(QJ_EARG, "null argument");
  if (option < 0 || option >= QJ_OPT_COUNT) return fail(QJ_EARG, "unknown option");
  *value = p->opt[option];
  return QJ_OK;
}

int qj_stage_count(void) { return kStageCount; }

const char* qj_stage_name(int stage) {
  return (stage >= 0 && stage < kStageCount) ? kStageNames[stage] : "";
}

int qj_timing_enable(qj_plan* p, int enable) {
  if (!p) return fail(QJ_EARG, "null plan");
  p->timing = enable ? 1 : 0;
  return QJ_OK;
}

int qj_timing_read(qj_plan* p, double* ms, long long* launches, int nstages) {
  if (!p || nstages < 0 || (nstages > 0 && (!ms || !launches)))
    return fail(QJ_EARG, "qj_timing_read: bad arguments");
  for (auto& r : p->recs) {
    float t = 0.f;
    hipError_t e = hipEventSynchronize(r.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
    if (e != hipSuccess) return hipfail(e, "qj_timing_read");
    p->stage_ms[r.stage] += t;
    p->stage_n[r.stage] += 1;
    p->pool.push_back(r.a);
    p->pool.push_back(r.b);
  }
  p->recs.clear();
  for (int k = 0; k < nstages && k < kStageCount; ++k) {
    ms[k] = p->stage_ms[k];
    launches[k] = p->stage_n[k];
    p->stage_ms[k] = 0.0;
    p->stage_n[k] = 0;
  }
  return QJ_OK;
}

size_t qj_workspace_size(const qj_plan* p) {
  if (!p) return 0;
  return 3 * align256(p->nhalf() * 2 * p->rsize) + align256(p->ncell() * p->rsize);
}

#define QJ_PRECISION(p, call_f, call_d) ((p)->dtype == QJ_C64 ? (call_f) : (call_d))

int qj_get_potential(qj_plan* p, const void* rho, double rho_bar, double four_pi_G,
                     double b_coeff, int dc_mode, void* V, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || !V || !work) return fail(QJ_EARG, "qj_get_potential: null buffer");
  hipStream_t s = (hipStream_t)stream;
  return QJ_PRECISION(p, get_potential_t<float>(p, rho, rho_bar, four_pi_G, b_coeff, dc_mode, V, work, s),
                      get_potential_t<double>(p, rho, rho_bar, four_pi_G, b_coeff, dc_mode, V, work, s));
}

int qj_acceleration_grid(qj_plan* p, const void* rho, double rho_bar, double four_pi_G,
                         int dc_mode, void* V, void* a_grid, double* diag, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || !a_grid || !work) return fail(QJ_EARG, "qj_acceleration_grid: null buffer");
  hipStream_t s = (hipStream_t)stream;
  return QJ_PRECISION(p, accel_grid_t<float>(p, rho, rho_bar, four_pi_G, dc_mode, V, a_grid, diag, work, s),
                      accel_grid_t<double>(p, rho, rho_bar, four_pi_G, dc_mode, V, a_grid, diag, work, s));
}

int qj_acceleration_grid_lp(qj_plan* p, const void* rho, double rho_bg, double four_pi_G,
                            double k_cut, void* V, void* a_grid, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || !a_grid || !work) return fail(QJ_EARG, "qj_acceleration_grid_lp: null buffer");
  if (!(k_cut > 0.0)) return fail(QJ_EARG, "qj_acceleration_grid_lp: k_cut must be positive");
  hipStream_t s = (hipStream_t)stream;
  const double kc2 = k_cut * k_cut + 1e-30;  // quantum-fluid-accretion.py:145
  return QJ_PRECISION(p, accel_grid_t<float>(p, rho, rho_bg, four_pi_G, 0, V, a_grid, nullptr,
                                             work, s, kc2),
                      accel_grid_t<double>(p, rho, rho_bg, four_pi_G, 0, V, a_grid, nullptr,
                                           work, s, kc2));
}

int qj_bondi_sink(qj_plan* p, void* rho, const void* vx, const void* vy, const void* vz,
                  const double* bh, double cs, double G, double dt, double r_acc_mult,
                  double lam, double fmax, double ann_lo, double ann_hi, double* out,
                  void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || !bh || !out) return fail(QJ_EARG, "qj_bondi_sink: null buffer");
  if ((vx || vy || vz) && !(vx && vy && vz))
    return fail(QJ_EARG, "qj_bondi_sink: give all three gas velocities or none");
  if (p->d.nx != p->d.ny || p->d.nx != p->d.nz || p->d.Lx != p->d.Ly || p->d.Lx != p->d.Lz)
    return fail(QJ_EUNSUPPORTED, "qj_bondi_sink: the sink needs a cubic grid and box");
  if (!(r_acc_mult > 0.0) || !(ann_lo > 0.0) || !(ann_hi > 0.0) || !(cs >= 0.0))
    return fail(QJ_EARG, "qj_bondi_sink: r_acc_mult, ann_lo, ann_hi must be positive");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = QJ_PRECISION(
      p,
      launch_bondi_sink<float>(p->d, rho, vx, vy, vz, bh, cs, G, dt, r_acc_mult, lam, fmax,
                               ann_lo, ann_hi, out, s),
      launch_bondi_sink<double>(p->d, rho, vx, vy, vz, bh, cs, G, dt, r_acc_mult, lam, fmax,
                                ann_lo, ann_hi, out, s));
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_bondi_sink");
}

int qj_kick(qj_plan* p, void* psi, const void* V, double coef, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!psi || !V) return fail(QJ_EARG, "qj_kick: null buffer");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = QJ_PRECISION(p, launch_kick<float>(psi, V, nullptr, nullptr, coef, p->ncell(), s),
                              launch_kick<double>(psi, V, nullptr, nullptr, coef, p->ncell(), s));
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_kick");
}

int qj_drift(qj_plan* p, void* psi, double dt, double m_per_hbar, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!psi) return fail(QJ_EARG, "qj_drift: null buffer");
  if (!(m_per_hbar != 0.0)) return fail(QJ_EARG, "qj_drift: m_per_hbar must be non-zero");
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == QJ_C64) {
    if (int rc = ensure_drift<float>(p, dt, m_per_hbar, s)) return rc;
    SpecTables<float> tb = tables<float>(p);
    for (int a = 2; a >= 0; --a)
      QJ_STAGE(kStageDriftX - a, launch_drift_axis<float>(p->d, tb, a, psi, nullptr, nullptr, nullptr, 0.0, 0, s), "drift");
  } else {
    if (int rc = ensure_drift<double>(p, dt, m_per_hbar, s)) return rc;
    SpecTables<double> tb = tables<double>(p);
    for (int a = 2; a >= 0; --a)
      QJ_STAGE(kStageDriftX - a, launch_drift_axis<double>(p->d, tb, a, psi, nullptr, nullptr, nullptr, 0.0, 0, s), "drift");
  }
  return QJ_OK;
}

int qj_cic_indices(qj_plan* p, const double* pos, long long np, int* i, int* ip1, double* w_i,
                   double* w_ip1, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (np < 0 || (np > 0 && (!pos || !i || !ip1 || !w_i || !w_ip1)))
    return fail(QJ_EARG, "qj_cic_indices: bad arguments");
  hipError_t e = launch_cic_indices<float>(p->d, pos, np, i, ip1, w_i, w_ip1, (hipStream_t)stream);
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_cic_indices");
}

int qj_bin_particles(qj_plan* p, const double* pos, const double* mass, double m_scalar,
                     long long np, void* rho, int accumulate, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || np < 0 || (np > 0 && !pos)) return fail(QJ_EARG, "qj_bin_particles: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (!accumulate)
    e = QJ_PRECISION(p, fill<float>(rho, p->ncell(), 0.0f, s), fill<double>(rho, p->ncell(), 0.0, s));
  if (e == hipSuccess)
    e = QJ_PRECISION(p, launch_deposit<float>(p->d, pos, mass, m_scalar, np, rho, s),
                     launch_deposit<double>(p->d, pos, mass, m_scalar, np, rho, s));
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_bin_particles");
}

int qj_interp(qj_plan* p, const double* pos, long long np, const void* a_grid, double* acc,
              void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (np < 0 || (np > 0 && (!pos || !a_grid || !acc))) return fail(QJ_EARG, "qj_interp: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (p->dtype == QJ_C64)
    e = launch_interp<float>(p->d, pos, np, a_grid, acc, s);
  else
    e = launch_interp<double>(p->d, pos, np, a_grid, acc, s);
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_interp");
}

int qj_get_acceleration(qj_plan* p, const double* pos, long long np, const void* rho,
                        double rho_bar, double four_pi_G, int dc_mode, double* acc, void* a_grid,
                        double* diag, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || !a_grid || !work || np < 0 || (np > 0 && (!pos || !acc)))
    return fail(QJ_EARG, "qj_get_acceleration: bad arguments");
  int rc = qj_acceleration_grid(p, rho, rho_bar, four_pi_G, dc_mode, nullptr, a_grid, diag, work, stream);
  if (rc) return rc;
  return qj_interp(p, pos, np, a_grid, acc, stream);
}

int qj_apply_grav_accel(qj_plan* p, const void* rho, void* vx, void* vy, void* vz, double dt,
                        double rho_bar, double four_pi_G, int dc_mode, void* a_grid, double* diag,
                        void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || !vx || !vy || !vz || !a_grid || !work)
    return fail(QJ_EARG, "qj_apply_grav_accel: null buffer");
  int rc = qj_acceleration_grid(p, rho, rho_bar, four_pi_G, dc_mode, nullptr, a_grid, diag, work, stream);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (p->dtype == QJ_C64)
    e = launch_gas_kick<float>(p->ncell(), vx, vy, vz, a_grid, dt, s);
  else
    e = launch_gas_kick<double>(p->ncell(), vx, vy, vz, a_grid, dt, s);
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_apply_grav_accel");
}

int qj_solve_hydro(qj_plan* p, const void* rho, const void* vx, const void* vy, const void* vz,
                   double dt, double cs, int flux, void* rho_o, void* vx_o, void* vy_o,
                   void* vz_o, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!rho || !vx || !vy || !vz || !rho_o || !vx_o || !vy_o || !vz_o)
    return fail(QJ_EARG, "qj_solve_hydro: null buffer");
  if (flux != QJ_FLUX_RUSANOV && flux != QJ_FLUX_HLL)
    return fail(QJ_EARG, "qj_solve_hydro: flux must be QJ_FLUX_RUSANOV or QJ_FLUX_HLL");
  if (p->d.nx != p->d.ny || p->d.nx != p->d.nz || p->d.Lx != p->d.Ly || p->d.Lx != p->d.Lz)
    return fail(QJ_EUNSUPPORTED, "qj_solve_hydro: the gas update needs a cubic grid and box");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = QJ_PRECISION(
      p,
      launch_hydro<float>(p->d, rho, vx, vy, vz, nullptr, 0.0, dt, cs, flux, rho_o,
                          vx_o, vy_o, vz_o, s),
      launch_hydro<double>(p->d, rho, vx, vy, vz, nullptr, 0.0, dt, cs, flux, rho_o,
                           vx_o, vy_o, vz_o, s));
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_solve_hydro");
}

}  // extern "C"

// ===========================================================================
// fused KDK stepping
// ===========================================================================
namespace {

// Columns of the fused star gather for a step of this state (0: the step
// writes the acceleration grid and the stars gather from it): stars with
// the tile-binned deposit, no gas, no tape (the adjoint reads the grid),
// and a Poisson x pass whose column tiles the deposit supports.
template <class T>
int gather_cols(const qj_plan* p, const qj_state* st, bool tape) {
  if (st->np <= 0 || !st->star_scratch || !p->opt[QJ_OPT_BINNED_DEPOSIT] || st->gas[0] || tape ||
      !p->opt[QJ_OPT_STAR_GATHER])
    return 0;
  return poisson_x_gather_columns(p->d, sizeof(T) == 4, p->opt[QJ_OPT_WIDE_SWEEPS]);
}

// y rows of the column tiles (QJ_OPT_GATHER_ROWS, 0 = automatic): two rows
// cut the stars' registrations by a quarter, but below 2 x 128-byte rows of
// columns they cut the Poisson x pass's V rows to half cache lines (+5 us at
// 256^3).  Measured A/B: 128^3 + 2^16 stars (one per 32 cells) 127.3 ->
// 123.9 us/step with two rows; 256^3 + 2^18 (one per 64 cells) 482.6 ->
// 487.0 us/step, and the 48-box ensemble with 2^14 stars per box 5.08 ->
// 5.33 ms/step.  Automatic: two rows where they keep whole lines or with at
// least one star per 32 cells.  With the fused deposit (fdep) a star
// registers in one row whatever the tile height, so two rows save nothing
// there and the automatic choice keeps one unless two keep whole lines
// (128^3 + 2^16 stars 117.1 -> 116.1 us/step with one row, 64^3 + 2^14
// 86.1 -> 85.1).
template <class T>
int gather_rows(const qj_plan* p, const qj_state* st, int cols, bool fdep) {
  const int o = p->opt[QJ_OPT_GATHER_ROWS];
  if (o == 1 || o == 2) return o;
  const int full = 128 / (int)sizeof(T);  // reals per 128-byte row
  if (cols / 2 >= full) return 2;
  if (fdep) return 1;
  return (st->np * 32 >= p->ncell()) ? 2 : 1;
}

// V and A3 of the current state (star deposit, source pass, Poisson chain);
// cols > 0: the fused star gather (V and the stars' corner values);
// fdep_mode > 0 (with cols): the deposit fused into the source pass on the
// one-row registration, as the steps run it (2: rho_s not written).
template <class T>
int initial_grid_t(qj_plan* p, qj_state* st, const qj_params* pr, const Work& w, hipStream_t s,
                   int cols = 0, int fdep_mode = 0) {
  SpecTables<T> tb = tables<T>(p);
  const bool stars = st->np > 0;
  const bool gas = st->gas[0] != nullptr;
  void* A3 = (stars || gas) ? plane<T>(st->grid, p, 1) : nullptr;
  void* brho = (pr->b_coeff != 0.0) ? st->brho : nullptr;
  const bool binned = st->star_scratch && p->opt[QJ_OPT_BINNED_DEPOSIT];
  const bool fused = stars && binned && cols && fdep_mode;
  StarGather ga{};
  StarDeposit sd{};
  if (stars && binned) {
    // tile-binned deposit of the current positions (overwrites rho_s), or
    // with the fused deposit the registrations only
    StarBins sb = carve_star_scratch(p->d, st->np, st->star_scratch, sizeof(T), cols,
                                     cols ? gather_rows<T>(p, st, cols, fused) : 1);
    sb.home = fused ? 1 : 0;
    if (cols) ga = StarGather{sb.offset, sb.sorted, sb.desc, sb.cv, sb.tg.lz, sb.home};
    if (fused)
      sd = StarDeposit{sb.offset, sb.sorted, sb.desc, sb.wts,
                       fdep_mode == 1 ? st->rho_s : nullptr, sb.tg.ly, sb.tg.lz, sb.tg.ntiles,
                       sb.home};
    QJ_CHECK(fill<int>(sb.count, (long long)sb.ntiles * kCountStride, 0, s), "bin reset");
    QJ_STAGE(kStageDeposit,
             launch_star_step_binned<T>(p->d, st->pos, st->vel, st->np, nullptr, 0.0, 0.0, 0.0, 0,
                                        sb, s),
             "bin");
    QJ_STAGE(kStageDeposit,
             launch_tile_deposit<T>(p->d, st->pos, st->mass, st->m_scalar, st->np, sb,
                                    fused ? nullptr : st->rho_s, s,
                                    p->opt[QJ_OPT_DETERMINISTIC_DEPOSIT]),
             "deposit");
  } else if (stars) {
    // the float-atomic deposit accumulates into a cleared plane
    QJ_CHECK(fill<T>(st->rho_s, p->ncell(), (T)0, s), "rho_s clear");
    QJ_STAGE(kStageDeposit,
             launch_deposit<T>(p->d, st->pos, st->mass, st->m_scalar, st->np, st->rho_s, s),
             "deposit");
  }
  const int clear = 0;  // both deposits leave rho_s as the source of this state
  if (stars) st->rho_valid = (fused && fdep_mode == 2) ? 0 : 1;
  void* src = stars ? (fused ? nullptr : st->rho_s)
                    : (gas ? (st->gas_swapped ? st->gas_tmp[0] : st->gas[0]) : nullptr);
  QJ_STAGE(kStageSourceX,
           launch_source_x<T>(p->d, tb, 0, st->psi, src, clear, brho, pr->b_coeff,
                              pr->rho_bar, w.H, p->partials, st->diag, s, nullptr,
                              fused ? &sd : nullptr),
           "source_x");
  if (int rc = poisson_tail<T>(p, w, (stars || gas) ? 1 : 0, pr->four_pi_G, pr->rho_bar,
                               pr->dc_mode, st->grid, A3, nullptr, st->diag, s, 0.0, nullptr,
                               ga.cv ? &ga : nullptr))
    return rc;
  st->grid_valid = ga.cv ? 2 : 1;
  return QJ_OK;
}

// The star density of the current positions into rho_s (qj_star_density):
// the stars registered in the deposit's own tiles and added with the
// fixed-point tile deposit (the same bits on every call), or with float
// atomics without a star scratch.  Only the registration part of the
// scratch is used, so the fused gather's corner values after it survive.
template <class T>
int star_density_t(qj_plan* p, qj_state* st, hipStream_t s) {
  if (st->np > 0) {
    if (!st->pos || !st->rho_s) return fail(QJ_EARG, "qj_star_density: stars need pos and rho_s");
    if (st->star_scratch && p->opt[QJ_OPT_BINNED_DEPOSIT]) {
      StarBins sb = carve_star_scratch(p->d, st->np, st->star_scratch);
      QJ_CHECK(fill<int>(sb.count, (long long)sb.ntiles * kCountStride, 0, s), "bin reset");
      QJ_CHECK(launch_star_step_binned<T>(p->d, st->pos, st->vel, st->np, nullptr, 0.0, 0.0, 0.0,
                                          0, sb, s),
               "bin");
      QJ_CHECK(launch_tile_deposit<T>(p->d, st->pos, st->mass, st->m_scalar, st->np, sb,
                                      st->rho_s, s, 1),
               "deposit");
    } else {
      QJ_CHECK(fill<T>(st->rho_s, p->ncell(), (T)0, s), "rho_s clear");
      QJ_CHECK(launch_deposit<T>(p->d, st->pos, st->mass, st->m_scalar, st->np, st->rho_s, s),
               "deposit");
    }
  }
  st->rho_valid = 1;
  return QJ_OK;
}

// dslot (the CFL stepper): this step's dt and kick time live in the device
// step slot; the drift tables are rebuilt from it each step and the owed
// kick stays on the device (the caller reads it back into st->pending).
// grid_out (one step only): the step's Poisson solve writes the new grid
// there and st->grid then points to it (the tape recorder uses this to
// write each step's grid straight into its tape).
template <class T>
// grid_out / psi_out (one step, the tape recorder): the Poisson solve writes
// its grid to grid_out, and psi after the drift goes to psi_out (the z sweep
// reads st->psi and writes psi_mid, the x sweep reads psi_mid and writes
// psi_out; fused x-source schedule only); st->grid / st->psi then point there.
// sig (gas steps of the CFL stepper, with dslot): the Poisson x pass also
// writes the CFL signal maxima of the state after the step (GasSignalFold).
int kdk_steps_t(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps, void* work,
                hipStream_t s, const double* kick_dt = nullptr, const double* drift_dt = nullptr,
                const double* dslot = nullptr, void* grid_out = nullptr, void* psi_mid = nullptr,
                void* psi_out = nullptr, GasSignalFold* sig = nullptr) {
  if (sig && (!dslot || !st->gas[0]))
    return fail(QJ_EARG, "kdk_steps_t: the signal fold needs the step slot and a gas state");
  if ((grid_out || psi_out) && nsteps != 1)
    return fail(QJ_EARG, "kdk_steps_t: grid_out / psi_out need one step");
  if (psi_out && (!psi_mid || !p->opt[QJ_OPT_FUSE_X_SOURCE]))
    return fail(QJ_EARG, "kdk_steps_t: psi_out needs psi_mid and the fused x-source schedule");
  Work w = carve(p, work);
  SpecTables<T> tb = tables<T>(p);
  if (!drift_dt && !dslot)
    if (int rc = ensure_drift<T>(p, pr->dt, pr->m_per_hbar, s)) return rc;
  const bool stars = st->np > 0;
  const bool gas = st->gas[0] != nullptr;
  const bool grad = stars || gas;
  void* V = st->grid;
  void* A3 = grad ? plane<T>(st->grid, p, 1) : nullptr;
  void* brho = (pr->b_coeff != 0.0) ? st->brho : nullptr;
  const double dt = pr->dt;
  auto live_gas = [&](int k) { return st->gas_swapped ? st->gas_tmp[k] : st->gas[k]; };
  auto spare_gas = [&](int k) { return st->gas_swapped ? st->gas[k] : st->gas_tmp[k]; };
  void* src = nullptr;
  const bool binned = st->star_scratch && p->opt[QJ_OPT_BINNED_DEPOSIT];
  const int clear = 0;  // both deposits leave rho_s as the source of this state
  const int cols = gather_cols<T>(p, st, grid_out != nullptr);
  // the star deposit fused into the source pass: on the gather's column
  // tiles, which are also the source pass's workgroup columns
  const int fd_opt = p->opt[QJ_OPT_FUSED_DEPOSIT];
  const bool fdep = cols && fd_opt && !p->opt[QJ_OPT_DETERMINISTIC_DEPOSIT] &&
                    cols == source_x_columns(p->d, sizeof(T) == 4);
  // the x schedule; the automatic one (2) splits the x drift from the source
  // pass where the star work runs on the side stream with the fused deposit
  // and the x pencils have 256 points or more: the split x drift then runs
  // beside the star work as well and the streams contend less (256^3 +
  // 2^18 stars 472.9 -> 469.1 us/step, 2^16 stars 423.2 -> 419.8; at 128^3
  // + 2^16 the fused pass is ahead, 115.8 against 121.6)
  const bool forks = stars && p->opt[QJ_OPT_SIDE_STREAM] && p->side && !g_recorder;
  const int fx_opt = p->opt[QJ_OPT_FUSE_X_SOURCE];
  const bool fuse_x = fx_opt == 1 || (fx_opt == 2 && !(forks && fdep && p->d.nx >= 256));
  if (psi_out && !fuse_x)
    return fail(QJ_EARG, "kdk_steps_t: psi_out needs the fused x-source schedule");
  // the corner values of the fused gather are served from the binned
  // deposit's scratch; the float-atomic star step reads the grid, so a state
  // that kept only corner values has its grid computed again
  if (st->grid_valid == 2 && !(stars && binned)) st->grid_valid = 0;

  for (int step = 0; step < nsteps; ++step) {
    if (!st->grid_valid) {
      if (int rc = initial_grid_t<T>(p, st, pr, w, s, cols, fdep ? fd_opt : 0)) return rc;
    }
    // the star registrations of this step (column tiles with the fused
    // gather); the kicks read the corner values when the previous Poisson
    // x pass left them instead of the grid
    StarBins sb{};
    if (stars && binned) {
      sb = carve_star_scratch(p->d, st->np, st->star_scratch, sizeof(T), cols,
                              cols ? gather_rows<T>(p, st, cols, fdep) : 1);
      // the fused deposit registers each star in its lower y row only
      sb.home = fdep ? 1 : 0;
    }
    const void* cv_in = (st->grid_valid == 2) ? sb.cv : nullptr;
    // per-step kick and drift times (a schedule, e.g. the comoving factors of
    // quantum-cosmo.py:170-175), else dt for both
    const double dk = kick_dt ? kick_dt[step] : dt;
    if (dslot) {
      const int n[3] = {p->d.nx, p->d.ny, p->d.nz};
      const double L[3] = {p->d.Lx, p->d.Ly, p->d.Lz};
      QJ_CHECK(launch_drift_tables3<T>(p->drift, n, L, 1.0 / (2.0 * pr->m_per_hbar), s, dslot),
               "drift tables");
      p->drift_key_dt = NAN;  // the tables now follow the device dt
    } else if (drift_dt) {
      if (int rc = ensure_drift<T>(p, drift_dt[step], pr->m_per_hbar, s)) return rc;
    }
    const double kick_t = st->pending + 0.5 * dk;  // owed half kick + this step's first
    // The star work of the step reads only the previous grid and the star
    // arrays, and the psi drift sweeps neither read nor write those, so with
    // the side-stream option it runs on the plan's second stream beside them;
    // the caller's stream waits for it before the Poisson source pass (which
    // reads rho_s).  The gas update stays on the caller's stream: its
    // workgroups fill every CU, so the sweeps beside it only waited for them
    // (the 256^3 gas step measured the same either way).  Recording (the
    // ensemble) keeps one stream.
    const bool fork = forks;
    auto side_work = [&](hipStream_t s) -> int {
      // stars: owed closing kick of the previous step and this step's opening
      // half kick (same grid, same positions), drift, deposit at the new positions
      if (stars && binned) {
        QJ_STAGE(kStageStarStep,
                 launch_star_step_binned<T>(p->d, st->pos, st->vel, st->np, A3, st->pending_star,
                                            0.5 * dt, dt, 1, sb, s, cv_in),
                 "star step");
        QJ_STAGE(kStageDeposit,
                 launch_tile_deposit<T>(p->d, st->pos, st->mass, st->m_scalar, st->np, sb,
                                        fdep ? nullptr : st->rho_s, s,
                                        p->opt[QJ_OPT_DETERMINISTIC_DEPOSIT]),
                 "deposit");
      } else if (stars) {
        // the float-atomic deposit accumulates into a cleared plane
        QJ_CHECK(fill<T>(st->rho_s, p->ncell(), (T)0, s), "rho_s clear");
        QJ_STAGE(kStageStarStep,
                 launch_star_step<T>(p->d, st->pos, st->vel, st->np, A3, st->pending_star,
                                     0.5 * dt, dt, st->mass, st->m_scalar, st->rho_s, s),
                 "star step");
      }
      // gas: owed + first half kick of the velocity, then the MUSCL update
      // over this step's drift time
      if (gas) {
        const double dd = drift_dt ? drift_dt[step] : dt;
        QJ_STAGE(kStageHydro,
                 launch_hydro<T>(p->d, live_gas(0), live_gas(1), live_gas(2), live_gas(3), A3,
                                 kick_t, dd, pr->cs, pr->flux, spare_gas(0), spare_gas(1),
                                 spare_gas(2), spare_gas(3), s, dslot),
                 "hydro");
        st->gas_swapped ^= 1;
      }
      return QJ_OK;
    };
    // psi: kick folded into the z drift sweep, then y and x (+ Poisson source)
    // With the star work beside it, the z sweep runs 1024 walking workgroups
    // (four pencil groups each at 256^3) instead of 512: the star kernels
    // then take CUs between its workgroups.  256^3 + 2^18 stars with the
    // fused deposit, us/step: one group per workgroup 469.3, 512 walking
    // 467-475, 768 469.5, 1024 464.0, 1536 462.5, 2048 474.8; with 2^16
    // stars 1024 and one group per workgroup take the same time.  Since the
    // one-row registration the star work is smaller and 512 are ahead again
    // at 256^3 (454.7 against 461.2 us/step, interleaved runs), while 1024
    // stay ahead below (128^3 + 2^16: 115.5 against 116.1)
    SpecTables<T> tbz = tb;
    if (fork && tbz.zblocks == 0 && p->ncell() < (1LL << 24)) tbz.zblocks = 1024;
    auto drift_zy = [&](hipStream_t s) -> int {
      void* mid = (psi_out && psi_mid != st->psi) ? psi_mid : nullptr;  // z sweep out of place
      QJ_STAGE(kStageDriftZKick,
               launch_drift_axis<T>(p->d, tbz, 2, st->psi, V, brho, st->v_ext,
                                    pr->m_per_hbar * (dslot ? 1.0 : kick_t), 1, s, 0, dslot, mid),
               "drift z");
      if (mid) st->psi = mid;
      QJ_STAGE(kStageDriftY,
               launch_drift_axis<T>(p->d, tb, 1, st->psi, nullptr, nullptr, nullptr, 0.0, 0, s),
               "drift y");
      return QJ_OK;
    };
    if (fork) {
      QJ_CHECK(hipEventRecord(p->ev_fork, s), "fork");
      QJ_CHECK(hipStreamWaitEvent(p->side, p->ev_fork, 0), "fork");
    }
    // (issuing the drift sweeps first measured the same)
    if (int rc = side_work(fork ? p->side : s)) return rc;
    if (int rc = drift_zy(s)) return rc;
    if (fork) QJ_CHECK(hipEventRecord(p->ev_join, p->side), "join");
    src = stars ? (fdep ? nullptr : st->rho_s) : (gas ? live_gas(0) : nullptr);
    StarDeposit sd{};
    if (fdep)
      sd = StarDeposit{sb.offset, sb.sorted, sb.desc, sb.wts, fd_opt == 1 ? st->rho_s : nullptr,
                       sb.tg.ly, sb.tg.lz, sb.tg.ntiles, sb.home};
    if (fork) QJ_CHECK(hipStreamWaitEvent(s, p->ev_join, 0), "join");
    if (fuse_x) {
      QJ_STAGE(kStageDriftXSource,
               launch_source_x<T>(p->d, tb, 1, st->psi, src, clear, brho, pr->b_coeff,
                                  pr->rho_bar, w.H, p->partials, st->diag, s, psi_out,
                                  fdep ? &sd : nullptr),
               "drift x + source");
      if (psi_out) st->psi = psi_out;
    } else {
      QJ_STAGE(kStageDriftX,
               launch_drift_axis<T>(p->d, tb, 0, st->psi, nullptr, nullptr, nullptr, 0.0, 0, s), "drift x");
      QJ_STAGE(kStageSourceX,
               launch_source_x<T>(p->d, tb, 0, st->psi, src, clear, brho, pr->b_coeff,
                                  pr->rho_bar, w.H, p->partials, st->diag, s, nullptr,
                                  fdep ? &sd : nullptr),
               "source x");
    }
    void* Vn = grid_out ? grid_out : V;
    void* A3n = grid_out ? (grad ? plane<T>(grid_out, p, 1) : nullptr) : A3;
    if (sig) {  // the velocities after this step's update, the kick it leaves owed
      for (int k = 0; k < 3; ++k) sig->v[k] = live_gas(k + 1);
      sig->owed = dslot + kSlotOwed;
      sig->cs = pr->cs;
      sig->part = p->sig_part;
      sig->cap = p->nblk_sig;
    }
    StarGather ga{};
    if (cols) ga = StarGather{sb.offset, sb.sorted, sb.desc, sb.cv, sb.tg.lz, sb.home};
    if (int rc = poisson_tail<T>(p, w, grad, pr->four_pi_G, pr->rho_bar, pr->dc_mode, Vn, A3n,
                                 nullptr, st->diag, s, 0.0, sig, cols ? &ga : nullptr))
      return rc;
    if (grid_out) st->grid = grid_out;
    st->grid_valid = cols ? 2 : 1;
    // the fused deposit without its rho_s write leaves the density uncomputed
    if (stars) st->rho_valid = (fdep && fd_opt == 2) ? 0 : 1;
    // the closing half kicks (psi, gas and stars) are owed to the next step
    if (!dslot) st->pending = 0.5 * dk;
    if (stars) st->pending_star = 0.5 * dt;
  }
  return QJ_OK;
}

// The CFL-limited gas stepper: before each step the device computes
// dt = min(dt_max, cfl dx / max(cs + |v|)) of the state the step starts
// from (quantum-fluid-accretion.py:696-701), and the step reads dt and its
// kick time from the device step slot, so steps follow each other without
// host round trips.  The host reads the chunk's dts and the owed kick once
// per kCflChunk steps.
template <class T>
int kdk_steps_cfl_t(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps, double cfl,
                    double dt_max, double* dts, void* work, hipStream_t s) {
  double* slot = p->partials + p->nblk_src + kDiagCount;
  double* dts_dev = slot + kSlotCount;
  const double dx = p->d.Lx / p->d.nx;
  QJ_CHECK(fill<double>(slot + kSlotOwed, 1, st->pending, s), "step slot");
  // The first step's signal comes from a pass over v and the grid; every
  // later step's was folded into the previous step's Poisson x pass.
  GasSignalFold sig{};
  bool folded = false;
  for (int k0 = 0; k0 < nsteps; k0 += kCflChunk) {
    const int m = nsteps - k0 < kCflChunk ? nsteps - k0 : kCflChunk;
    for (int k = 0; k < m; ++k) {
      if (folded) {
        // the previous step's Poisson x pass left the signal maxima
        QJ_CHECK(launch_cfl_finish(p->sig_part, sig.nblk, cfl, dx, dt_max, slot, dts_dev, k, s),
                 "cfl");
      } else {
        void* const* g = st->gas_swapped ? st->gas_tmp : st->gas;
        // the owed kick is zero on a state whose grid is not computed yet
        const void* A3 = st->grid_valid ? plane<T>(st->grid, p, 1) : nullptr;
        QJ_CHECK(launch_gas_cfl<T>(p->ncell(), g[1], g[2], g[3], A3, 0.0, pr->cs, dx, cfl,
                                   dt_max, slot, dts_dev, k, p->partials, p->nblk_src, s),
                 "cfl");
      }
      if (int rc = kdk_steps_t<T>(p, st, pr, 1, work, s, nullptr, nullptr, slot, nullptr,
                                  nullptr, nullptr, &sig))
        return rc;
      folded = true;
    }
    double h[kCflChunk + 1];
    hipError_t e = hipMemcpyAsync(h, dts_dev, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h + m, slot + kSlotOwed, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hipfail(e, "qj_kdk_steps_cfl: read-back");
    st->pending = h[m];
    for (int k = 0; k < m; ++k) {
      if (dts) dts[k0 + k] = h[k];
      if (!(h[k] > 0.0) || !std::isfinite(h[k]))
        return fail(QJ_EARG, "qj_kdk_steps_cfl: non-finite CFL time step (non-finite gas velocity?)");
    }
  }
  return QJ_OK;
}

template <class T>
int kdk_finalize_t(qj_plan* p, qj_state* st, const qj_params* pr, hipStream_t s) {
  void* V = st->grid;
  void* A3 = plane<T>(st->grid, p, 1);
  if (st->pending != 0.0) {
    void* brho = (pr->b_coeff != 0.0) ? st->brho : nullptr;
    QJ_STAGE(kStageKick,
             launch_kick<T>(st->psi, V, brho, st->v_ext, pr->m_per_hbar * st->pending, p->ncell(), s), "kick");
    if (st->gas[0]) {
      void* g1 = st->gas_swapped ? st->gas_tmp[1] : st->gas[1];
      void* g2 = st->gas_swapped ? st->gas_tmp[2] : st->gas[2];
      void* g3 = st->gas_swapped ? st->gas_tmp[3] : st->gas[3];
      QJ_STAGE(kStageGasKick, launch_gas_kick<T>(p->ncell(), g1, g2, g3, A3, st->pending, s),
               "gas kick");
    }
    st->pending = 0.0;
  }
  if (st->pending_star != 0.0 && st->np > 0) {
    // the corner values of the fused gather, when the last Poisson x pass
    // left those instead of the grid
    const void* cv = (st->grid_valid == 2)
                         ? carve_star_scratch(p->d, st->np, st->star_scratch, sizeof(T)).cv
                         : nullptr;
    QJ_CHECK(fill<double>(st->diag + kDiagDtS, 1, __builtin_inf(), s), "dt_s reset");
    QJ_STAGE(kStageStarKick,
             launch_star_kick<T>(p->d, st->pos, st->vel, st->np, A3, st->pending_star, st->diag, s,
                                 cv),
             "star kick");
    st->pending_star = 0.0;
  }
  return QJ_OK;
}

// ---------------------------------------------------------------------------
// recording and adjoint of the (unmerged) KDK step, quantum-stars.py:295-330
// ---------------------------------------------------------------------------
size_t grid_bytes(const qj_plan* p) { return (1 + kAccComp) * (size_t)p->ncell() * p->rsize; }

template <class T>
int kdk_record_t(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps, qj_tape_step* tape,
                 void* grid0, void* work, hipStream_t s) {
  Work w = carve(p, work);
  if (int rc = ensure_drift<T>(p, pr->dt, pr->m_per_hbar, s)) return rc;
  const size_t cb = (size_t)p->ncell() * 2 * sizeof(T);
  const size_t xb = (size_t)st->np * 3 * sizeof(double);
  // the tape holds whole grids: a state whose stars kept only their corner
  // values (grid_valid 2) has its grid recomputed first
  if (st->grid_valid != 1)
    if (int rc = initial_grid_t<T>(p, st, pr, w, s)) return rc;
  if (grid0) QJ_CHECK(copy_async(grid0, st->grid, grid_bytes(p), s), "tape grid0");
  // each step's Poisson solve writes its grid straight into the tape (the
  // next step reads it there); the state's own grid gets the last one back.
  // With the fused x-source schedule psi after the drift goes straight into
  // the tape as well: the x sweep writes psi2 of the step, the next step's z
  // sweep reads it there and writes the state's own psi.
  void* const own_grid = st->grid;
  void* const own_psi = st->psi;
  const bool psi_direct = p->opt[QJ_OPT_FUSE_X_SOURCE] != 0;
  int rc = QJ_OK;
  // (every exit below restores st->grid and st->psi: the tape buffers are
  // the caller's)
  auto copy = [&](void* dst, const void* src, size_t bytes, const char* what) {
    const hipError_t e = copy_async(dst, src, bytes, s);
    if (e != hipSuccess) rc = hipfail(e, what);
  };
  for (int k = 0; k < nsteps && rc == QJ_OK; ++k) {
    qj_tape_step& t = tape[k];
    if (t.psi0) copy(t.psi0, st->psi, cb, "tape psi0");
    if (rc == QJ_OK && st->np > 0 && t.x0) copy(t.x0, st->pos, xb, "tape x0");
    // merged steps, as the forward runs them: the closing half kick of step k
    // is owed to (and merged into) the opening kick of step k + 1
    if (rc == QJ_OK)
      rc = kdk_steps_t<T>(p, st, pr, 1, work, s, nullptr, nullptr, nullptr, t.grid1,
                          psi_direct ? own_psi : nullptr, psi_direct ? t.psi2 : nullptr);
    if (rc == QJ_OK && !psi_direct) copy(t.psi2, st->psi, cb, "tape psi2");
    if (rc == QJ_OK && st->np > 0) copy(t.x1, st->pos, xb, "tape x1");
    // the tile registrations the step's binned deposit made of x1
    if (rc == QJ_OK && t.bins && st->np > 0 && st->star_scratch && p->opt[QJ_OPT_BINNED_DEPOSIT])
      copy(t.bins, st->star_scratch, star_scratch_bytes(p->d, st->np), "tape bins");
  }
  if (st->grid != own_grid) {
    const hipError_t e = copy_async(own_grid, st->grid, grid_bytes(p), s);
    st->grid = own_grid;
    if (e != hipSuccess && rc == QJ_OK) rc = hipfail(e, "tape: grid back into the state");
  }
  if (st->psi != own_psi) {
    const hipError_t e = copy_async(own_psi, st->psi, cb, s);
    st->psi = own_psi;
    if (e != hipSuccess && rc == QJ_OK) rc = hipfail(e, "tape: psi back into the state");
  }
  if (rc) return rc;
  return kdk_finalize_t<T>(p, st, pr, s);
}

template <class T>
int kdk_adjoint_t(qj_plan* p, const qj_state* st, const qj_params* pr, int nsteps,
                  const qj_tape_step* tape, const void* grid0, qj_adjoint* a, void* work,
                  hipStream_t s) {
  Work w = carve(p, work);
  if (int rc = ensure_drift<T>(p, pr->dt, pr->m_per_hbar, s)) return rc;
  SpecTables<T> tb = tables<T>(p);
  const long long n = p->ncell();
  const double dt = pr->dt, h = 0.5 * dt, c = pr->m_per_hbar * h;
  const bool stars = st->np > 0;
  T* Vbar = (T*)a->scratch;
  T* abar3 = Vbar + n;
  T* rhobar = abar3 + kAccComp * n;
  const bool binned = st->star_scratch && p->opt[QJ_OPT_BINNED_DEPOSIT];
  // one adjoint half step: the kick of psi (V, psi_ref) and of the stars
  // (A3, pos) with kick time hh (phase coefficient cc), the density and
  // Poisson adjoints, the deposit adjoint.  kick_done: the psi kick's adjoint
  // (psi_bar phase and Vbar) was applied by the z sweep of the drift adjoint
  // before it; defer_rho: the density adjoint is left to the x sweep of the
  // drift adjoint after it (both fusions keep the formulas of the separate
  // kernels)
  auto half = [&](const void* grid, const void* psi_ref, const double* pos, double dt_xv,
                  double hh, double cc, const void* bins, bool kick_done,
                  bool defer_rho) -> int {
    const void* V = grid;
    const void* A3 = (const char*)grid + (size_t)n * sizeof(T);
    if (stars && binned) {
      // the star kicks' adjoint deposited tile-binned (writes every cell of
      // abar3), with the forward's taped registrations of pos when given
      StarBins sb = carve_star_scratch(p->d, st->np, bins ? const_cast<void*>(bins)
                                                          : st->star_scratch);
      QJ_STAGE(kStageAdjParticles,
               launch_interp_adj_binned<T>(p->d, pos, st->np, A3, hh, a->v_bar, dt_xv, a->x_bar,
                                           abar3, sb, bins != nullptr, s),
               "interp adjoint");
    } else {
      QJ_CHECK(fill<T>(abar3, kAccComp * n, (T)0, s), "abar clear");
      if (stars)
        QJ_STAGE(kStageAdjParticles,
                 launch_interp_adj<T>(p->d, pos, st->np, A3, hh, a->v_bar, dt_xv, a->x_bar, abar3,
                                      s),
                 "interp adjoint");
    }
    if (!kick_done)
      QJ_STAGE(kStageAdjKick,
               launch_kick_adj<T>(a->psi_bar, psi_ref, V, st->v_ext, cc, Vbar, n, s),
               "kick adjoint");
    QJ_STAGE(kStageAdjPoisson,
             launch_adj_poisson<T>(p->d, tb, Vbar, abar3, pr->four_pi_G, pr->dc_mode, w.H, w.Uz,
                                   w.Uy, rhobar, s),
             "adjoint Poisson");
    if (!defer_rho)
      QJ_STAGE(kStageAdjRho, launch_rho_adj<T>(a->psi_bar, psi_ref, rhobar, n, s), "rho adjoint");
    if (stars)
      QJ_STAGE(kStageAdjParticles,
               launch_deposit_adj<T>(p->d, pos, st->mass, st->m_scalar, st->np, rhobar, a->m_bar,
                                     a->x_bar, s),
               "deposit adjoint");
    return QJ_OK;
  };
  // Reverse sweep.  The closing kick of step k-1 and the opening kick of step
  // k act with one potential (grid1 of step k-1) at one set of positions, so
  // their adjoints are one half step with doubled kick time, taken at the
  // input of the pair (psi2 and x1 of step k-1): Im(conj(psi_bar) psi) is
  // the same before and after the common phase, and both star kicks see the
  // same v_bar.  One adjoint Poisson solve per step, as the forward merges
  // the kicks.
  const void* rho_ref = nullptr;  // psi_ref of the deferred density adjoint
  for (int k = nsteps - 1; k >= 0; --k) {
    const qj_tape_step& t = tape[k];
    // closing half step of the last step
    if (k == nsteps - 1) {
      if (int rc = half(t.grid1, t.psi2, t.x1, 0.0, h, c, t.bins, false, true)) return rc;
      rho_ref = t.psi2;
    }
    // the half step after the drift: opening half of step k with the closing
    // half of step k-1, or the opening half step of the first step (grid0 at
    // x0, psi0)
    const void* ngrid = k > 0 ? tape[k - 1].grid1 : grid0;
    const void* nref = k > 0 ? tape[k - 1].psi2 : t.psi0;
    const double ncc = k > 0 ? 2.0 * c : c;
    // drift: psi_bar <- F^H psi_bar (conjugate phase tables = drift of -dt),
    // x sweep first (with the deferred density adjoint), z sweep last (with
    // the next half step's kick adjoint); pos1 = pos0 + dt vel1 is folded
    // into the next interp adjoint (dt_xv)
    QJ_STAGE(kStageAdjDrift,
             launch_drift_adj_axis<T>(p->d, tb, 0, a->psi_bar, rhobar, rho_ref, nullptr, nullptr,
                                      0.0, nullptr, nullptr, s),
             "drift adjoint");
    QJ_STAGE(kStageAdjDrift,
             launch_drift_axis<T>(p->d, tb, 1, a->psi_bar, nullptr, nullptr, nullptr, 0.0, 0, s, 1),
             "drift adjoint");
    QJ_STAGE(kStageAdjDrift,
             launch_drift_adj_axis<T>(p->d, tb, 2, a->psi_bar, nullptr, nullptr, ngrid, st->v_ext,
                                      ncc, nref, Vbar, s),
             "drift adjoint");
    if (k > 0) {
      const qj_tape_step& u = tape[k - 1];
      if (int rc = half(u.grid1, u.psi2, u.x1, dt, 2.0 * h, 2.0 * c, u.bins, true, true)) return rc;
      rho_ref = u.psi2;
    } else {
      if (int rc = half(grid0, t.psi0, t.x0, dt, h, c, nullptr, true, false)) return rc;
    }
  }
  return QJ_OK;
}

}  // namespace

extern "C" {

int qj_kdk_steps(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps, void* work,
                 void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!st || !pr || !work || nsteps < 0) return fail(QJ_EARG, "qj_kdk_steps: bad arguments");
  if (pr->flux != QJ_FLUX_RUSANOV && pr->flux != QJ_FLUX_HLL)
    return fail(QJ_EARG, "qj_kdk_steps: params.flux must be QJ_FLUX_RUSANOV or QJ_FLUX_HLL");
  if (!st->psi || !st->grid || !st->diag) return fail(QJ_EARG, "qj_kdk_steps: null state buffer");
  if (st->np > 0 && (!st->pos || !st->vel || !st->rho_s))
    return fail(QJ_EARG, "qj_kdk_steps: stars need pos, vel and rho_s");
  if (st->gas[0] && (!st->gas[1] || !st->gas[2] || !st->gas[3] || !st->gas_tmp[0] ||
                     !st->gas_tmp[1] || !st->gas_tmp[2] || !st->gas_tmp[3]))
    return fail(QJ_EARG, "qj_kdk_steps: gas needs rho, vx, vy, vz and their second buffers");
  if (st->gas[0] && st->np > 0)
    return fail(QJ_EUNSUPPORTED, "qj_kdk_steps: stars and gas in one state");
  if (st->gas[0] && (p->d.nx != p->d.ny || p->d.nx != p->d.nz || p->d.Lx != p->d.Ly ||
                     p->d.Lx != p->d.Lz))
    return fail(QJ_EUNSUPPORTED, "qj_kdk_steps: the gas update needs a cubic grid and box");
  if (pr->b_coeff != 0.0 && !st->brho)
    return fail(QJ_EARG, "qj_kdk_steps: self-interaction needs the brho buffer");
  if (!(pr->m_per_hbar != 0.0)) return fail(QJ_EARG, "qj_kdk_steps: m_per_hbar must be non-zero");
  if (st->v_ext && (st->np > 0 || st->gas[0]))
    return fail(QJ_EUNSUPPORTED,
                "qj_kdk_steps: the external potential acts on psi only (no stars or gas)");
  hipStream_t s = (hipStream_t)stream;
  return p->dtype == QJ_C64 ? kdk_steps_t<float>(p, st, pr, nsteps, work, s)
                            : kdk_steps_t<double>(p, st, pr, nsteps, work, s);
}

int qj_kdk_steps_noalloc(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps,
                         void* stream) {
  if (int rc = check_plan(p)) return rc;
  const size_t need = qj_workspace_size(p);
  if (need > (size_t)QJ_NOALLOC_SCRATCH_BYTES)
    return fail(QJ_EUNSUPPORTED,
                "qj_kdk_steps_noalloc: the plan's workspace (" + std::to_string(need) +
                    " bytes) exceeds the " + std::to_string((size_t)QJ_NOALLOC_SCRATCH_BYTES) +
                    "-byte static workspace; use qj_kdk_steps with a caller workspace");
  if (p->timing)
    return fail(QJ_EARG, "qj_kdk_steps_noalloc: stage timing creates events; disable it first");
  void* work = nullptr;
  if (int rc = scratch_address(p->device, &work)) return rc;
  // the workspace holds only spectra that each step writes before reading
  // them, so nothing needs clearing between calls
  return qj_kdk_steps(p, st, pr, nsteps, work, stream);
}

int qj_kdk_steps_sched(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps,
                       const double* kick_dt, const double* drift_dt, void* work, void* stream) {
  if (nsteps > 0 && (!kick_dt || !drift_dt))
    return fail(QJ_EARG, "qj_kdk_steps_sched: null schedule");
  if (st && st->np > 0)
    return fail(QJ_EUNSUPPORTED, "qj_kdk_steps_sched: psi and gas states (no stars)");
  if (int rc = qj_kdk_steps(p, st, pr, 0, work, stream)) return rc;  // argument checks
  hipStream_t s = (hipStream_t)stream;
  return p->dtype == QJ_C64
             ? kdk_steps_t<float>(p, st, pr, nsteps, work, s, kick_dt, drift_dt)
             : kdk_steps_t<double>(p, st, pr, nsteps, work, s, kick_dt, drift_dt);
}

int qj_gas_cfl_dt(qj_plan* p, const qj_state* st, const qj_params* pr, double cfl, double dt_max,
                  double* out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!st || !pr || !out || !(cfl > 0.0) || !(dt_max > 0.0))
    return fail(QJ_EARG, "qj_gas_cfl_dt: bad arguments");
  if (!st->gas[0] || !st->gas_tmp[0]) return fail(QJ_EARG, "qj_gas_cfl_dt: the state has no gas");
  const int k0 = st->gas_swapped;
  void* const* g = k0 ? st->gas_tmp : st->gas;
  // the owed half kick of the last step is part of the reference's state
  const double h = (st->grid_valid && st->pending != 0.0) ? st->pending : 0.0;
  const void* A3 = h != 0.0 ? (p->dtype == QJ_C64 ? plane<float>(st->grid, p, 1)
                                                  : plane<double>(st->grid, p, 1))
                            : nullptr;
  const double dx = p->d.Lx / p->d.nx;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = QJ_PRECISION(
      p,
      launch_gas_cfl<float>(p->ncell(), g[1], g[2], g[3], A3, h, pr->cs, dx, cfl, dt_max, out,
                            nullptr, 0, p->partials, p->nblk_src, s),
      launch_gas_cfl<double>(p->ncell(), g[1], g[2], g[3], A3, h, pr->cs, dx, cfl, dt_max, out,
                             nullptr, 0, p->partials, p->nblk_src, s));
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_gas_cfl_dt");
}

int qj_kdk_steps_cfl(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps, double cfl,
                     double dt_max, double* dts, void* work, void* stream) {
  if (nsteps < 0 || !(cfl > 0.0) || !(dt_max > 0.0))
    return fail(QJ_EARG, "qj_kdk_steps_cfl: bad arguments");
  if (int rc = qj_kdk_steps_sched(p, st, pr, 0, nullptr, nullptr, work, stream)) return rc;
  if (!st->gas[0] || !st->gas_tmp[0]) return fail(QJ_EARG, "qj_kdk_steps_cfl: the state has no gas");
  hipStream_t s = (hipStream_t)stream;
  return p->dtype == QJ_C64 ? kdk_steps_cfl_t<float>(p, st, pr, nsteps, cfl, dt_max, dts, work, s)
                            : kdk_steps_cfl_t<double>(p, st, pr, nsteps, cfl, dt_max, dts, work, s);
}

int qj_bh_accretion(qj_plan* p, void* psi, const double* pos, const double* vel, double* m_bh,
                    long long n_bh, double G, double c, double sigma, double m_per_hbar,
                    double dt, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!psi || !pos || !vel || !m_bh || n_bh < 1)
    return fail(QJ_EARG, "qj_bh_accretion: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = QJ_PRECISION(
      p,
      launch_bh_accretion<float>(p->d, psi, pos, vel, m_bh, n_bh, G, c, sigma, m_per_hbar, dt, s),
      launch_bh_accretion<double>(p->d, psi, pos, vel, m_bh, n_bh, G, c, sigma, m_per_hbar, dt,
                                  s));
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_bh_accretion");
}

size_t qj_star_scratch_size(const qj_plan* p, long long np) {
  if (!p || np < 0) return 0;
  return star_scratch_total_bytes(p->d, np, (int)p->rsize);
}

size_t qj_adjoint_scratch_size(const qj_plan* p) {
  return p ? (2 + kAccComp) * (size_t)p->ncell() * p->rsize : 0;
}

int qj_kdk_record(qj_plan* p, qj_state* st, const qj_params* pr, int nsteps, qj_tape_step* tape,
                  void* grid0, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!st || !pr || !work || nsteps < 0 || (nsteps > 0 && !tape))
    return fail(QJ_EARG, "qj_kdk_record: bad arguments");
  if (st->gas[0]) return fail(QJ_EUNSUPPORTED, "qj_kdk_record: the gas step has no adjoint");
  if (pr->b_coeff != 0.0)
    return fail(QJ_EUNSUPPORTED, "qj_kdk_record: self-interaction has no adjoint");
  if (st->pending != 0.0 || st->pending_star != 0.0)
    return fail(QJ_EARG, "qj_kdk_record: finalize the state first (owed half kick)");
  for (int k = 0; k < nsteps; ++k)
    if ((k == 0 && !tape[k].psi0) || !tape[k].psi2 || !tape[k].grid1 ||
        (st->np > 0 && ((k == 0 && !tape[k].x0) || !tape[k].x1)))
      return fail(QJ_EARG, "qj_kdk_record: null tape buffer");
  hipStream_t s = (hipStream_t)stream;
  return p->dtype == QJ_C64 ? kdk_record_t<float>(p, st, pr, nsteps, tape, grid0, work, s)
                            : kdk_record_t<double>(p, st, pr, nsteps, tape, grid0, work, s);
}

int qj_kdk_adjoint(qj_plan* p, const qj_state* st, const qj_params* pr, int nsteps,
                   const qj_tape_step* tape, const void* grid0, qj_adjoint* adj, void* work,
                   void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!st || !pr || !adj || !work || nsteps < 0 || (nsteps > 0 && (!tape || !grid0)))
    return fail(QJ_EARG, "qj_kdk_adjoint: bad arguments");
  if (!adj->psi_bar || !adj->scratch ||
      (st->np > 0 && (!adj->x_bar || !adj->v_bar || !adj->m_bar)))
    return fail(QJ_EARG, "qj_kdk_adjoint: null adjoint buffer");
  if (st->gas[0] || pr->b_coeff != 0.0)
    return fail(QJ_EUNSUPPORTED, "qj_kdk_adjoint: gas and self-interaction have no adjoint");
  if (nsteps > 0 && !tape[0].psi0) return fail(QJ_EARG, "qj_kdk_adjoint: null psi0 of step 0");
  if (nsteps > 0 && st->np > 0 && !tape[0].x0)
    return fail(QJ_EARG, "qj_kdk_adjoint: null x0 of step 0");
  hipStream_t s = (hipStream_t)stream;
  return p->dtype == QJ_C64
             ? kdk_adjoint_t<float>(p, st, pr, nsteps, tape, grid0, adj, work, s)
             : kdk_adjoint_t<double>(p, st, pr, nsteps, tape, grid0, adj, work, s);
}

int qj_deposit_adjoint(qj_plan* p, const double* pos, const double* mass, double m_scalar,
                       long long np, const void* rhobar, double* m_bar, double* x_bar,
                       void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (np < 0 || (np > 0 && (!pos || !rhobar || !m_bar || !x_bar)))
    return fail(QJ_EARG, "qj_deposit_adjoint: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e =
      QJ_PRECISION(p, launch_deposit_adj<float>(p->d, pos, mass, m_scalar, np, rhobar, m_bar, x_bar, s),
                   launch_deposit_adj<double>(p->d, pos, mass, m_scalar, np, rhobar, m_bar, x_bar, s));
  return e == hipSuccess ? QJ_OK : hipfail(e, "qj_deposit_adjoint");
}

int qj_star_density(qj_plan* p, qj_state* st, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!st) return fail(QJ_EARG, "qj_star_density: null state");
  hipStream_t s = (hipStream_t)stream;
  return p->dtype == QJ_C64 ? star_density_t<float>(p, st, s) : star_density_t<double>(p, st, s);
}

int qj_kdk_finalize(qj_plan* p, qj_state* st, const qj_params* pr, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!st || !pr) return fail(QJ_EARG, "qj_kdk_finalize: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  return p->dtype == QJ_C64 ? kdk_finalize_t<float>(p, st, pr, s)
                            : kdk_finalize_t<double>(p, st, pr, s);
}

}  // extern "C"
