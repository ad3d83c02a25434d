// solver.cpp -- implementation of the rism_* C ABI (include/rism3d_amd.h):
// handle set-up, device buffers, the iterate driver and the thermodynamic
// reductions.  The driver restates src-cuda/iterate.cu:6-90 (restart,
// initial estimate, charge-up, step loop, divergence and step-limit
// handling) on top of the fused kernels in k_iterate.hip.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "rism_host.h"
#include "rism_kernels.h"

using namespace rism;

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

bool file_exists(const char *p) {
  struct stat st;
  return p && p[0] && stat(p, &st) == 0;
}

// factor n into radices {16, 8, 4, 2, 3, 5}: as few power-of-two stages as
// possible with their exponents spread evenly, then the odd radices
// radices of an n-point plan: the power-of-two part in stages of at most
// 2^maxbits (balanced bit counts), then the threes and fives
bool make_radices(int n, std::vector<int> &rad, int maxbits = 4) {
  rad.clear();
  int e = 0, m = n;
  while (m % 2 == 0) { m /= 2; ++e; }
  if (e > 0) {
    const int ns = (e + maxbits - 1) / maxbits;
    for (int s = 0; s < ns; ++s) {
      const int b = e / ns + (s < e % ns ? 1 : 0);
      rad.push_back(1 << b);
    }
  }
  while (m % 3 == 0) { rad.push_back(3); m /= 3; }
  while (m % 5 == 0) { rad.push_back(5); m /= 5; }
  return m == 1 && (int)rad.size() <= kMaxStages;
}

int ept_class(long elems) {
  const long per = (elems + kThreads - 1) / kThreads;
  if (per <= 4) return 4;
  if (per <= 8) return 8;
  if (per <= 16) return 16;
  if (per <= 32) return 32;
  return -1;
}

int largest_divisor_le(long n, long cap) {
  for (long d = std::min(n, cap); d >= 1; --d)
    if (n % d == 0) return (int)d;
  return 1;
}

}  // namespace

// Everything the iteration kernels need in one field precision R: FFT plans
// with twiddle tables, launch geometry, the closure potential u, the
// long-range terms, the chi spline table and the spectrum work array.
struct PrecSet {
  bool ready = false;
  AxisPlan plan[3];
  AxisPlan plan_xr;  // x axis, radices <= 5 (closure pass, FP64 residual pass), same tables as plan[0]
  LaunchGeom geom{};
  void *u = nullptr, *fr = nullptr, *fkh = nullptr, *chi = nullptr, *Z = nullptr, *tables = nullptr;
};

template <typename R> constexpr int pidx() { return sizeof(R) == sizeof(double) ? 0 : 1; }

struct rism_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  int nx = 0, ny = 0, nz = 0;
  long N = 0;
  double box[3] = {0, 0, 0}, dr[3] = {0, 0, 0}, dv = 0, volume = 0;
  int V = 0, P = 0;
  std::vector<double> rho, qv, sigv, epsv;
  double temper = 0, xt = 0, dk = 0;
  int ntab = 0;
  std::vector<double> chi;  // [ntab][V][V]
  int natu = 0;
  int natu_cap = 1;  // atoms the solute-sized device arrays hold
  std::vector<double> usig, ueps, uq, uxyz;
  rism_opts opts{};
  int hm = 2;               // compiled history class >= nhist
  PrecSet ps[2];            // [0] FP64 (always), [1] FP32 (mixed precision)
  bool tuned = false;       // rism_autotune has chosen the tile shapes
  // device memory
  std::vector<DevBuf> bufs;
  // batched solves: the handle's grid-sized buffers are carved, in the same
  // order for every solute of a batch, from its slice of one arena
  // (arena_cap bytes at arena), so each field of solute s sits s * stride
  // bytes after solute 0's; arena_measure records the slice size instead.
  // Solute-sized buffers (atom_alloc) never come from the arena.
  char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0, arena_need = 0;
  bool arena_measure = false, atom_alloc = false;
  // a batch member after the first: the plans, their tables and the chi
  // spline tables are the first member's (same grid and solvent), not copies
  const rism_handle *tmpl = nullptr;
  bool borrowed_ctx = false;  // stream and pinned control block belong to the batch
  void *scratch = nullptr;  // one-off reductions (see handle_scratch())
  size_t scratch_bytes = 0;
  double *ulj = nullptr, *de = nullptr, *fr = nullptr, *sr = nullptr;
  double *sel = nullptr, *seq = nullptr;  // [V][N] SE atom sums from the potentials pass
  double2 *fkh = nullptr, *phase = nullptr;
  // mixed precision state (k_mixed_impl.h): reference T of t_s and A of
  // the closure argument, the change dl, the refresh residual R_ref, the closure
  // pass's B, the increment history and the residual ring
  double *mT = nullptr;
  float *mA = nullptr;  // A = T - u, FP32
  float *mdl = nullptr, *mrref = nullptr, *mbq = nullptr, *mhx = nullptr, *mring = nullptr;
  double *ts = nullptr;
  void *res = nullptr, *hist_t = nullptr, *hist_r = nullptr;  // FP64, or FP32 in the mixed path
  size_t qsz = 8;           // bytes per residual / history element
  double *partials = nullptr, *rms_hist = nullptr, *gfield = nullptr;
  Ctl *ctl = nullptr;
  Ctl *ctl_host = nullptr;  // pinned
  double4 *atoms = nullptr;
  double2 *lj = nullptr;
  double *sigh = nullptr, *d_qv = nullptr, *d_rho = nullptr;
  // adjoint (allocated on the first rism_adjoint_grad)
  double *adj_f = nullptr;       // gp, alpha, beta, b, lam, res, ubar: 7 x [V][N]
  double *adj_hist = nullptr;    // hist_l, hist_r: 2 x [nhist][V][N]
  double *adj_n = nullptr;       // debar, frbar: 2 x [N]
  double2 *adj_Z = nullptr, *adj_Y = nullptr;
  double2 *adj_ljd = nullptr;   // [V][natu]: sized by the solute (adj_ljd_cap atoms)
  int adj_ljd_cap = 0;
  double *adj_rms = nullptr;
  Ctl *adj_ctl = nullptr;
  // state
  bool potentials = false, solved = false, evaluated = false;
  double cf = 1.0;
  int refreshes = 0;        // mixed: FP64 refresh steps of the last solve
  bool mx_live = false;     // mixed: A, dl (not ts) hold the current state
  // replayable chunks of iterations: [precision][0]
  hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int gchunk[2][2] = {{0, 0}, {0, 0}};
  double gcf[2][2] = {{0, 0}, {0, 0}};  // charge-up factor the graph was captured with
  bool graph_ok = true;  // false after a capture failed on this handle
  int steps = 0;
  double rms = 0.0;
  double t_setup = 0, t_solve = 0, t_thermo = 0;
  std::vector<double> rms_log;
  // per-kernel HIP-event timing (rism_profile)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_next = 0;
  std::vector<std::pair<int, size_t>> ev_marks;  // (kernel id, index of start event)
  double kt_ms[RISM_NKERNEL] = {0};
  long kt_n[RISM_NKERNEL] = {0};
};

#define HIP_OK(...)                                                                        \
  do {                                                                                     \
    hipError_t _e = (__VA_ARGS__);                                                         \
    if (_e != hipSuccess) {                                                                \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #__VA_ARGS__); \
      return RISM_ERR_HIP;                                                                 \
    }                                                                                      \
  } while (0)

namespace {

// ---- per-kernel timing: an event pair around each launch on the stream ----
int prof_mark(rism_handle *h, int kernel) {
  if (!h->prof) return RISM_OK;
  while (h->ev_pool.size() < h->ev_next + 2) {
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    h->ev_pool.push_back(e);
  }
  HIP_OK(hipEventRecord(h->ev_pool[h->ev_next], h->stream));
  h->ev_marks.push_back({kernel, h->ev_next});
  h->ev_next += 2;
  return RISM_OK;
}

int prof_done(rism_handle *h) {
  if (!h->prof) return RISM_OK;
  HIP_OK(hipEventRecord(h->ev_pool[h->ev_marks.back().second + 1], h->stream));
  return RISM_OK;
}

int prof_collect(rism_handle *h) {
  if (!h->prof || h->ev_marks.empty()) return RISM_OK;
  HIP_OK(hipStreamSynchronize(h->stream));
  for (auto &m : h->ev_marks) {
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, h->ev_pool[m.second], h->ev_pool[m.second + 1]));
    h->kt_ms[m.first] += ms;
    h->kt_n[m.first] += 1;
  }
  h->ev_marks.clear();
  h->ev_next = 0;
  return RISM_OK;
}

#define PROF(h, id, ...)                                  \
  do {                                                    \
    if (int _rc = prof_mark((h), (id))) return _rc;       \
    HIP_OK(__VA_ARGS__);                                  \
    if (int _rc = prof_done(h)) return _rc;               \
  } while (0)

// Device allocation and release (which synchronise the device) never
// overlap the recording of a HIP graph on another handle's stream: handles
// used from several host threads at once take this lock around both.
static std::mutex g_device_mu;

// blocking copy ordered on the handle's own stream (never on the legacy
// stream, which would tie this handle to others, e.g. one being captured)
hipError_t hcopy(rism_handle *h, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, h->stream);
  return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
}

// Device buffers of destroyed handles are kept for reuse, per device and
// exact size, up to RISM_POOL_BYTES in all (default 8 GiB; 0 disables):
// hipFree synchronises the whole device, and a batch that creates one handle
// per solute would otherwise stall every other worker's stream at each
// release.  An allocation that fails first returns the cache to the device.
// Guarded by g_device_mu.
struct CachedBuf {
  int device;
  size_t bytes;
  void *p;
};
std::vector<CachedBuf> g_cache;
size_t g_cache_bytes = 0;

long env_long(const char *name, long dflt);

size_t cache_cap() {
  static const size_t cap = (size_t)std::max(0L, env_long("RISM_POOL_BYTES", 8L << 30));
  return cap;
}

// A destroyed handle's stream (synchronised, graphs dropped) and pinned
// control block are kept for the next handle on the same device, up to 64:
// creating and destroying them also synchronises.  Guarded by g_device_mu.
struct SpareCtx {
  int device;
  hipStream_t stream;
  Ctl *ctl_host;
};
std::vector<SpareCtx> g_spare;
constexpr size_t kMaxSpare = 64;

void cache_flush_locked(int device) {
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (size_t i = 0; i < g_cache.size();) {
    if (g_cache[i].device == device) {
      (void)hipSetDevice(device);
      (void)hipFree(g_cache[i].p);
      g_cache_bytes -= g_cache[i].bytes;
      g_cache[i] = g_cache.back();
      g_cache.pop_back();
    } else {
      ++i;
    }
  }
  (void)hipSetDevice(cur);
}

int dalloc(rism_handle *h, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (!h->atom_alloc) {
    const size_t b = (bytes + 255) & ~(size_t)255;
    if (h->arena) {  // a batch member: the next piece of its arena slice (not in bufs)
      if (h->arena_used + b > h->arena_cap) {
        set_error("batch arena slice exhausted");
        return RISM_ERR_NOMEM;
      }
      *p = h->arena + h->arena_used;
      h->arena_used += b;
      return RISM_OK;
    }
    if (h->arena_measure) h->arena_need += b;
  }
  hipError_t e = hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_device_mu);
    *p = nullptr;
    for (size_t i = 0; i < g_cache.size(); ++i)
      if (g_cache[i].device == h->device && g_cache[i].bytes == bytes) {
        *p = g_cache[i].p;
        g_cache_bytes -= bytes;
        g_cache[i] = g_cache.back();
        g_cache.pop_back();
        break;
      }
    if (!*p) {
      e = hipMalloc(p, bytes);
      if (e != hipSuccess && !g_cache.empty()) {
        (void)hipGetLastError();
        cache_flush_locked(h->device);
        e = hipMalloc(p, bytes);
      }
    }
  }
  if (e != hipSuccess) {
    set_error("device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    *p = nullptr;
    return RISM_ERR_NOMEM;
  }
  h->bufs.push_back({*p, bytes});
  return RISM_OK;
}

// a buffer to the cache, or back to the device when the cache is full
void release_locked(int device, const DevBuf &b) {
  if (g_cache_bytes + b.bytes <= cache_cap()) {
    g_cache.push_back({device, b.bytes, b.p});
    g_cache_bytes += b.bytes;
  } else {
    (void)hipFree(b.p);
  }
}

// return one of the handle's buffers once its stream's queued work is done
void dfree(rism_handle *h, void *p) {
  if (!p) return;
  (void)hipStreamSynchronize(h->stream);
  std::lock_guard<std::mutex> lk(g_device_mu);
  for (size_t i = 0; i < h->bufs.size(); ++i)
    if (h->bufs[i].p == p) {
      release_locked(h->device, h->bufs[i]);
      h->bufs[i] = h->bufs.back();
      h->bufs.pop_back();
      return;
    }
}

template <typename T> int dalloc_t(rism_handle *h, T **p, size_t count) {
  void *q = nullptr;
  int rc = dalloc(h, &q, count * sizeof(T));
  *p = (T *)q;
  return rc;
}

// per-handle scratch for the one-off reductions (grad, euv, qv): grown on
// demand through the buffer cache, never hipMalloc/hipFree per call
int handle_scratch(rism_handle *h, size_t bytes, void **p) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (h->scratch_bytes < bytes) {
    dfree(h, h->scratch);
    h->scratch = nullptr;
    h->scratch_bytes = 0;
    if (int rc = dalloc(h, &h->scratch, bytes)) return rc;
    h->scratch_bytes = bytes;
  }
  *p = h->scratch;
  return RISM_OK;
}
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// per-axis plan + twiddle tables in the iteration precision
template <typename R> int build_plans(rism_handle *h) {
  using T2 = typename C2<R>::T;
  PrecSet &ps = h->ps[pidx<R>()];
  const int ns[3] = {h->nx, h->ny, h->nz};
  size_t tot = 0;
  for (int a = 0; a < 3; ++a) tot += 2 * (size_t)ns[a];
  T2 *dt = nullptr;
  if (int rc = dalloc_t(h, &dt, tot)) return rc;
  std::vector<T2> host(tot);
  // Stockham stage parameters of an n-point plan with radices up to 2^maxbits
  auto fill = [&](AxisPlan &p, int n, int maxbits) -> int {
    std::vector<int> rad;
    if (!make_radices(n, rad, maxbits)) {
      set_error("unsupported grid length " + std::to_string(n));
      return RISM_ERR_ARG;
    }
    std::memset(&p, 0, sizeof p);
    p.n = n;
    p.nst = (int)rad.size();
    int nn = n, s = 1;
    for (int k = 0; k < p.nst; ++k) {
      p.rad[k] = rad[k];
      p.s[k] = s;
      p.m[k] = nn / rad[k];
      p.tstep[k] = n / nn;
      p.ls2[k] = (s & (s - 1)) == 0 ? __builtin_ctz(s) : -1;
      const int nb = n / rad[k];
      p.lb2[k] = (nb & (nb - 1)) == 0 ? __builtin_ctz(nb) : -1;
      nn /= rad[k];
      s *= rad[k];
    }
    p.fixsign = ((n / 2) % 2) ? -1 : 1;
    return RISM_OK;
  };
  size_t off = 0;
  for (int a = 0; a < 3; ++a) {
    const int n = ns[a];
    AxisPlan &p = ps.plan[a];
    if (int rc = fill(p, n, 4)) return rc;
    for (int k = 0; k < n; ++k) {
      const double ang = 2.0 * M_PI * (double)k / (double)n;
      host[off + k].x = (R)std::cos(ang);
      host[off + k].y = (R)-std::sin(ang);
      const double ph = M_PI * (double)(k - n / 2) / (double)n;
      const double sg = (k & 1) ? -1.0 : 1.0;
      host[off + n + k].x = (R)(sg * std::cos(ph));
      host[off + n + k].y = (R)(-sg * std::sin(ph));
    }
    p.tw = dt + off;
    p.pre = dt + off + n;
    off += 2 * (size_t)n;
  }
  // the closure pass's and the FP64 residual pass's x plan: radix-4 stages
  // at most (those kernels compile no radix-8/16 stage), on the x axis's
  // tables
  if (int rc = fill(ps.plan_xr, h->nx, 2)) return rc;
  ps.plan_xr.tw = ps.plan[0].tw;
  ps.plan_xr.pre = ps.plan[0].pre;
  HIP_OK(hcopy(h, dt, host.data(), tot * sizeof(T2), hipMemcpyHostToDevice));
  ps.tables = dt;
  return RISM_OK;
}

// per-workgroup partial sums of the residual pass are sized for this many
// workgroups (the autotuner's largest candidate)
constexpr int kMaxXBlocks = 8192;

long env_long(const char *name, long dflt) {
  const char *s = std::getenv(name);
  return (s && *s) ? std::atol(s) : dflt;
}

// Tile shapes: the residual pass holds Lx whole x lines (about 2048
// elements in FP32, 1024 in FP64) and the closure/update pass Lc lines
// (about 1024 elements: its
// many operand streams per point want more workgroups in flight), the y
// pass Ty consecutive-x columns, the z pass two mirrored tiles of Tz columns
// per pack.  Every workgroup's LDS (padded lines) stays within RISM_LDS_CAP
// bytes (64 KiB by default).  RISM_LX / RISM_LC / RISM_TY / RISM_TZ override
// the choices (tuning sweeps); `ovr` = {Lx, Ty, Tz, xblocks, Lc}, when given,
// takes their place (the autotuner's candidates, 0 = keep the default).
int choose_geometry(rism_handle *h, LaunchGeom &g, size_t rsz, const int *ovr = nullptr) {
  const int nx = h->nx, ny = h->ny, nz = h->nz, P = h->P;
  const long c2 = 2 * (long)rsz;
  // LDS per workgroup: 64 KiB by default (RISM_LDS_CAP raises or lowers it),
  // never above what the device grants one workgroup
  int lds_max = 65536;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) != hipSuccess)
    lds_max = 65536;
  const long cap = std::min<long>(env_long("RISM_LDS_CAP", 65536), lds_max);
  // residual pass: 1024 elements per workgroup (four x lines at 256: its
  // registers then allow four waves per SIMD, and the FP32 pass holds all
  // of a group's operands in flight during the transform)
  const long xe = env_long("RISM_X_ELEMS", 1024);
  const long lsx = line_stride(nx), lsy = line_stride(ny), lsz = line_stride(nz);
  const long ce = env_long("RISM_C_ELEMS", 1024);
  g.Lx = largest_divisor_le((long)ny * nz,
                            std::max(1L, std::min(xe / ((long)P * nx), (cap - 2L * nx * c2) / ((long)P * lsx * c2))));
  g.Lc = largest_divisor_le((long)ny * nz,
                            std::max(1L, std::min(ce / ((long)P * nx), (cap - 2L * nx * c2) / ((long)P * lsx * c2))));
  // the y and z passes also stage their twiddle and phase tables (2n complex) in LDS
  g.Ty = largest_divisor_le(nx, std::max(1L, std::min(8192L / ny, (cap - 2L * ny * c2) / (lsy * c2))));
  g.Tz = largest_divisor_le(nx, std::max(1L, std::min(8192L / (2L * P * nz), (cap - 2L * nz * c2) / (2L * P * lsz * c2))));
  const long lx = ovr ? ovr[0] : env_long("RISM_LX", 0), ty = ovr ? ovr[1] : env_long("RISM_TY", 0),
             tz = ovr ? ovr[2] : env_long("RISM_TZ", 0);
  const long lc = (ovr && ovr[4] > 0) ? ovr[4] : env_long("RISM_LC", 0);
  if (lx > 0 && ((long)ny * nz) % lx == 0) g.Lx = (int)lx;
  if (lc > 0 && ((long)ny * nz) % lc == 0) g.Lc = (int)lc;
  if (ty > 0 && nx % ty == 0) g.Ty = (int)ty;
  if (tz > 0 && nx % tz == 0) g.Tz = (int)tz;
  g.ept_x = ept_class((long)P * g.Lx * nx);
  g.xgroups = (int)((long)ny * nz / g.Lx);
  const long xb = (ovr && ovr[3] > 0) ? ovr[3] : env_long("RISM_XBLOCKS", 1024);
  g.xblocks = (int)std::min<long>(g.xgroups, std::min<long>(xb, kMaxXBlocks));
  g.lds_x = ((size_t)P * g.Lx * lsx + 2 * (size_t)nx) * c2;
  g.ept_c = ept_class((long)P * g.Lc * nx);
  g.cgroups = (int)((long)ny * nz / g.Lc);
  g.lds_c = ((size_t)P * g.Lc * lsx + 2 * (size_t)nx) * c2;
  g.ept_y = ept_class((long)g.Ty * ny);
  g.lds_y = ((size_t)g.Ty * lsy + 2 * (size_t)ny) * c2;
  g.ept_z = ept_class(2L * P * g.Tz * nz);
  g.lds_z = ((size_t)2 * P * g.Tz * lsz + 2 * (size_t)nz) * c2;
  if (g.ept_x < 0 || g.ept_x > 16 || g.ept_c < 0 || g.ept_c > 16 || g.ept_y < 0 || g.ept_z < 0 ||
      (long)g.lds_x > cap || (long)g.lds_c > cap || (long)g.lds_y > cap || (long)g.lds_z > cap) {
    set_error("grid too large for the single-workgroup line transforms");
    return RISM_ERR_ARG;
  }
  return RISM_OK;
}

template <typename R> IterArgs<R, double> make_args(rism_handle *h, bool force_mx = false) {
  IterArgs<R, double> a;
  std::memset(&a, 0, sizeof a);
  const PrecSet &ps = h->ps[pidx<R>()];
  a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.N = h->N;
  a.V = h->V; a.P = h->P;
  a.ts = h->ts;
  using Q = typename HistT<R>::T;
  if (sizeof(Q) == h->qsz) {  // the FP64 passes of the mixed path never touch them
    a.res = (Q *)h->res;
    a.hist_t = (Q *)h->hist_t;
    a.hist_r = (Q *)h->hist_r;
  }
  a.rms_hist = h->rms_hist;
  a.u = (const R *)ps.u;
  a.fr = (const R *)ps.fr;
  a.Z = (typename C2<R>::T *)ps.Z;
  a.fkh = (const typename C2<R>::T *)ps.fkh;
  a.chi = (const typename R4T<R>::T *)ps.chi;
  a.ntab = h->ntab;
  a.dk = h->dk;
  for (int k = 0; k < 3; ++k) a.box[k] = h->box[k];
  for (int v = 0; v < h->V; ++v) {
    a.qcf[v] = h->qv[v] * h->cf;
    a.rho_inv[v] = 1.0 / h->rho[v];
  }
  a.dv = h->dv;
  a.ivol = 1.0 / h->volume;
  a.closure = h->opts.closure;
  a.ctl = h->ctl;
  a.partials = h->partials;
  a.nv = 1 + h->hm + h->hm * (h->hm + 1) / 2;
  a.px = ps.plan[0]; a.py = ps.plan[1]; a.pz = ps.plan[2];
  a.pxr = ps.plan_xr;
  a.Lx = ps.geom.Lx; a.Lc = ps.geom.Lc; a.Ty = ps.geom.Ty; a.Tz = ps.geom.Tz;
  a.xblocks = ps.geom.xblocks;
  a.store_h = 0;
  a.use_fk = 1;
  if (h->opts.precision == RISM_PREC_MIXED) {
    if (pidx<R>() == 1 || force_mx) a.ts = h->mT;  // the mixed chains' T
    a.A = h->mA;
    a.dl = h->mdl;
    a.rref = h->mrref;
    a.bq = h->mbq;
    a.hx = h->mhx;
    a.rring = h->mring;
  }
  return a;
}

// One step of the mixed path: gate 1 the FP32 chain on differences to the
// reference, gate 2 the FP64 refresh chain (k_mixed_impl.h).  Each launch
// runs only at a step of its chain (the solve kernel decides the refreshes),
// so both chains can sit in one stream or graph.
int launch_mx_iteration(rism_handle *h, int gate, bool solve = true) {
  if (gate == 2) {
    IterArgs<double, double> a = make_args<double>(h, true);
    const LaunchGeom &g = h->ps[0].geom;
    a.gate = 2;
    a.use_fk = 1;
    PROF(h, 10, launch_mx_closure_fx<double>(a, g, h->hm, h->stream));
    PROF(h, 10, launch_fft_y<double, double>(a, g, false, h->stream));
    PROF(h, 10, launch_fft_z_oz<double, double>(a, g, h->stream));
    PROF(h, 10, launch_fft_y<double, double>(a, g, true, h->stream));
    PROF(h, 10, launch_mx_resid<double>(a, g, h->hm, h->stream));
    if (solve)
      PROF(h, 10, launch_anderson_solve(h->ctl, h->partials, g.xblocks, a.nv, h->hm, h->rms_hist, h->stream, 2));
    return RISM_OK;
  }
  IterArgs<float, double> a = make_args<float>(h);
  const LaunchGeom &g = h->ps[1].geom;
  a.gate = 1;
  a.use_fk = 0;  // the change of c_s has no long-range part
  PROF(h, 0, launch_mx_closure_fx<float>(a, g, h->hm, h->stream));
  PROF(h, 1, launch_fft_y<float, double>(a, g, false, h->stream));
  PROF(h, 2, launch_fft_z_oz<float, double>(a, g, h->stream));
  PROF(h, 3, launch_fft_y<float, double>(a, g, true, h->stream));
  PROF(h, 4, launch_mx_resid<float>(a, g, h->hm, h->stream));
  if (solve) PROF(h, 5, launch_anderson_solve(h->ctl, h->partials, g.xblocks, a.nv, h->hm, h->rms_hist, h->stream, 1));
  return RISM_OK;
}

template <typename R> int launch_iteration(rism_handle *h, bool store_h, bool solve,
                                          bool h_only = false) {
  if constexpr (pidx<R>() == 1) {
    return launch_mx_iteration(h, 1, solve);  // the mixed path's FP32 chain
  } else {
  IterArgs<R, double> a = make_args<R>(h);
  const LaunchGeom &g = h->ps[pidx<R>()].geom;
  a.store_h = store_h ? 1 : 0;
  a.href_only = h_only ? 1 : 0;
  PROF(h, 0, launch_closure_fx<R, double>(a, g, true, h->hm, h->stream));
  PROF(h, 1, launch_fft_y<R, double>(a, g, false, h->stream));
  PROF(h, 2, launch_fft_z_oz<R, double>(a, g, h->stream));
  PROF(h, 3, launch_fft_y<R, double>(a, g, true, h->stream));
  PROF(h, 4, launch_fx_resid<R, double>(a, g, h->hm, h->stream));
  if (solve)
    PROF(h, 5, launch_anderson_solve(h->ctl, h->partials, g.xblocks, a.nv, h->hm, h->rms_hist,
                                     h->stream));
  return RISM_OK;
  }
}

// the pending Anderson update alone (no closure / FFT): in the mixed path
// through the closure pass of the chain whose step is current
int launch_update_only(rism_handle *h) {
  if (h->opts.precision == RISM_PREC_MIXED) {
    if (h->ctl_host->refresh_iter == h->ctl_host->cur) {
      IterArgs<double, double> a = make_args<double>(h, true);
      a.gate = 2;
      HIP_OK(launch_mx_closure_fx<double>(a, h->ps[0].geom, h->hm, h->stream));
    } else {
      IterArgs<float, double> a = make_args<float>(h);
      a.gate = 1;
      HIP_OK(launch_mx_closure_fx<float>(a, h->ps[1].geom, h->hm, h->stream));
    }
    return RISM_OK;
  }
  IterArgs<double, double> a = make_args<double>(h);
  HIP_OK(launch_closure_fx<double, double>(a, h->ps[0].geom, true, h->hm, h->stream));
  return RISM_OK;
}

// mixed path: the state (A, dl) from t_s (dir 0, a new reference) or t_s
// from the state (dir 1), with the current stage's FP64 closure potential
int mx_convert(rism_handle *h, int dir) {
  if (h->opts.precision != RISM_PREC_MIXED) return RISM_OK;
  HIP_OK(launch_mx_convert(h->ts, h->mT, h->mA, h->mdl, (const double *)h->ps[0].u, (long)h->V * h->N, dir,
                           h->stream));
  return RISM_OK;
}

// the closure exponent uses u = beta u_LJ + q_v cf (de - fr) for this stage
template <typename R> int stage_potential(rism_handle *h) {
  std::vector<double> q(h->V);
  for (int v = 0; v < h->V; ++v) q[v] = h->qv[v] * h->cf;
  HIP_OK(launch_combine_u<R>(h->ulj, h->sr, q.data(), h->V, h->N, (R *)h->ps[pidx<R>()].u, h->stream));
  return RISM_OK;
}

// (the mixed path's FP32 chain works on differences and needs no potential)
int stage_potentials(rism_handle *h) { return stage_potential<double>(h); }

int put_ctl(rism_handle *h) {
  HIP_OK(hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  return RISM_OK;
}

int get_ctl(rism_handle *h) {
  HIP_OK(hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  return RISM_OK;
}

// full t = ts + q_v cf fr written to host [V][N]
int fetch_t(rism_handle *h, double *out) {
  const size_t VN = (size_t)h->V * h->N;
  if (h->mx_live)
    if (int rc = mx_convert(h, 1)) return rc;
  std::vector<double> fr(h->N);
  HIP_OK(hcopy(h, out, h->ts, VN * sizeof(double), hipMemcpyDeviceToHost));
  HIP_OK(hcopy(h, fr.data(), h->fr, h->N * sizeof(double), hipMemcpyDeviceToHost));
  for (int v = 0; v < h->V; ++v) {
    const double q = h->qv[v] * h->cf;
    for (long i = 0; i < h->N; ++i) out[(size_t)v * h->N + i] += q * fr[i];
  }
  return RISM_OK;
}

int store_t(rism_handle *h, const double *t) {
  const size_t VN = (size_t)h->V * h->N;
  std::vector<double> fr(h->N), ts(VN);
  HIP_OK(hcopy(h, fr.data(), h->fr, h->N * sizeof(double), hipMemcpyDeviceToHost));
  for (int v = 0; v < h->V; ++v) {
    const double q = h->qv[v] * h->cf;
    for (long i = 0; i < h->N; ++i) ts[(size_t)v * h->N + i] = t[(size_t)v * h->N + i] - q * fr[i];
  }
  HIP_OK(hcopy(h, h->ts, ts.data(), VN * sizeof(double), hipMemcpyHostToDevice));
  h->mx_live = false;
  return RISM_OK;
}

int write_tuv_file(rism_handle *h, const char *path) {
  // write_tuv.cu:7-33: header "nx ny nz natv", then one value per line,
  // default stream precision (6 significant digits)
  const size_t VN = (size_t)h->V * h->N;
  std::vector<double> t(VN);
  if (int rc = fetch_t(h, t.data())) return rc;
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("cannot write ") + path);
    return RISM_ERR_IO;
  }
  std::fprintf(f, "%d %d %d %d\n", h->nx, h->ny, h->nz, h->V);
  for (size_t i = 0; i < VN; ++i) std::fprintf(f, "%g\n", t[i]);
  std::fclose(f);
  return RISM_OK;
}

int read_tuv_file(rism_handle *h, const char *path) {
  FILE *f = std::fopen(path, "r");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return RISM_ERR_IO;
  }
  int a, b, c, d;
  if (std::fscanf(f, "%d %d %d %d", &a, &b, &c, &d) != 4) {
    std::fclose(f);
    set_error("malformed .tuv header");
    return RISM_ERR_FORMAT;
  }
  if (a != h->nx || b != h->ny || c != h->nz || d != h->V) {
    // read_tuv.cu:14-20 warns and keeps reading; a short file is an error here
    std::printf(" RXRISM:  actual   NGr = %ld   NatV = %d\n           saved   NGr = %ld   NatV = %d\n",
                h->N, h->V, (long)a * b * c, d);
  }
  const size_t VN = (size_t)h->V * h->N;
  std::vector<double> t(VN);
  for (size_t i = 0; i < VN; ++i)
    if (std::fscanf(f, "%lf", &t[i]) != 1) {
      std::fclose(f);
      set_error(".tuv file is shorter than the grid");
      return RISM_ERR_FORMAT;
    }
  std::fclose(f);
  return store_t(h, t.data());
}

static void drop_graphs(rism_handle *h) {
  for (int p = 0; p < 2; ++p)
    for (int d = 0; d < 2; ++d) {
      if (h->gexec[p][d]) (void)hipGraphExecDestroy(h->gexec[p][d]);
      h->gexec[p][d] = nullptr;
      h->gchunk[p][d] = 0;
    }
}

// n iterations as one HIP graph: captured on first use for the current
// precision and refinement mode, then replayed, also by later solves on the
// handle.  The launches of an iteration do not depend on its index (the
// kernels read it from the control block), and they hold no memset or copy
// nodes; of the handle's state they carry only q_v cf (captured again when
// cf differs) and the tile shapes (the autotuner drops the graphs).
template <typename R> int launch_chunk_graph(rism_handle *h, int n) {
  hipGraphExec_t &ex = h->gexec[pidx<R>()][0];
  int &gn = h->gchunk[pidx<R>()][0];
  if (!h->graph_ok) {
    for (int k = 0; k < n; ++k)
      if (int rc = launch_iteration<R>(h, false, true)) return rc;
    return RISM_OK;
  }
  double &gc = h->gcf[pidx<R>()][0];
  if (!ex || gn != n || gc != h->cf) {
    if (ex) (void)hipGraphExecDestroy(ex);
    ex = nullptr;
    // relaxed mode: other host threads (other handles) keep allocating and
    // copying while this stream records
    std::unique_lock<std::mutex> lk(g_device_mu);
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed);
    int rc = RISM_OK;
    hipGraph_t g = nullptr;
    if (e == hipSuccess) {
      for (int k = 0; k < n && rc == RISM_OK; ++k) rc = launch_iteration<R>(h, false, true);
      e = hipStreamEndCapture(h->stream, &g);
    }
    if (e == hipSuccess && rc == RISM_OK) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    lk.unlock();
    if (e != hipSuccess || rc != RISM_OK) {  // no graph: launch this handle's chunks eagerly
      (void)hipGetLastError();
      if (ex) (void)hipGraphExecDestroy(ex);
      ex = nullptr;
      h->graph_ok = false;
      for (int k = 0; k < n; ++k)
        if (int r2 = launch_iteration<R>(h, false, true)) return r2;
      return RISM_OK;
    }
    gn = n;
    gc = h->cf;
  }
  HIP_OK(hipGraphLaunch(ex, h->stream));
  return RISM_OK;
}

template <typename R> int solve_impl(rism_handle *h, const char *tuv_path, int *steps_out, double *rms_out) {
  const double t0 = now_s();
  const int V = h->V;
  const long N = h->N;
  const size_t VN = (size_t)V * N;
  constexpr bool kMixed = pidx<R>() == 1;
  const int H = h->opts.nhist;
  int cu = h->opts.chargeup;
  double cuf = 0.0;
  const bool verbose = h->opts.verbose != 0;
  const bool trace = env_long("RISM_TRACE", 0) > 0;  // refresh steps of the mixed path on stderr
  h->evaluated = false;
  // history and state start from zero (anderson.cu allocates without clearing;
  // unfilled slots are never read here)
  float *hx32 = kMixed ? h->mhx : nullptr;
  if (kMixed) {
    HIP_OK(hipMemsetAsync(h->mhx, 0, (size_t)H * VN * sizeof(float), h->stream));
    HIP_OK(hipMemsetAsync(h->mring, 0, (size_t)(H + 1) * VN * sizeof(float), h->stream));
  } else {
    HIP_OK(hipMemsetAsync(h->hist_t, 0, (size_t)H * VN * h->qsz, h->stream));
    HIP_OK(hipMemsetAsync(h->hist_r, 0, (size_t)H * VN * h->qsz, h->stream));
    HIP_OK(hipMemsetAsync(h->res, 0, VN * h->qsz, h->stream));
  }
  h->mx_live = false;
  if (file_exists(tuv_path)) {  // iterate.cu:25-33
    h->cf = 1.0;
    if (int rc = read_tuv_file(h, tuv_path)) return rc;
    cu = 0;
  } else if (cu == 0) {  // initialize_tuv.cu: t = q_v fr, i.e. ts = 0 at cf = 1
    if (verbose) std::printf("synthesizing initial estimate for Tuv ...\n");
    h->cf = 1.0;
    HIP_OK(hipMemsetAsync(h->ts, 0, VN * sizeof(double), h->stream));
  } else {  // charge-up from t = 0, cf = 0
    h->cf = 0.0;
    cuf = 1.0 / cu;
    HIP_OK(hipMemsetAsync(h->ts, 0, VN * sizeof(double), h->stream));
  }
  Ctl &c = *h->ctl_host;
  std::memset(&c, 0, sizeof c);
  c.count = h->opts.count;
  c.nhist = H;
  c.maxstep = h->opts.maxstep;
  c.tol = h->opts.tol;
  c.m = h->opts.m;
  c.mp = h->opts.mp;
  c.npts = (double)VN;
  c.upd_iter = -1;
  c.upd_hold = -1;
  c.refresh_iter = -1;
  if (kMixed) {  // residual ring: slot H holds the current R, slot s history s
    c.ring = 1;
    c.rcur = H;
    for (int k = 0; k < H; ++k) c.rmap[k] = k;
  }
  int status = RISM_SOLVE_CONVERGED;
  h->rms_log.clear();
  h->refreshes = 0;
  // iteration chunks replay as HIP graphs unless timing or tracing every
  // step (RISM_GRAPHS=0 turns them off)
  const char *genv = std::getenv("RISM_GRAPHS");
  const bool graphs = !(genv && genv[0] == '0') && !h->opts.verbose && !h->prof && h->opts.ksave <= 0;
  // iterations enqueued per control-block read; 1 when timing kernels so no
  // gated (empty) launches enter the averages; RISM_CHUNK (1..16) overrides
  // the default (tuning).  At 256^3 (mixed) 8 measured 0.7 ms per job faster
  // than 4: fewer host round trips against the gated launches a chunk
  // queues beyond a refresh request or the stop
  const long chunk_env = env_long("RISM_CHUNK", 0);
  const int chunk = (h->opts.verbose || h->prof)
                        ? 1
                        : (chunk_env >= 1 && chunk_env <= kRmsRecent ? (int)chunk_env : (N >= (1L << 21) ? 8 : 16));
  int printed = 0;
  for (int stage = 0; stage <= cu; ++stage) {
    if (stage > 0) {  // add_tuv.cu: t += q_v cuf fr; ts absorbs any clamp remainder
      if (kMixed)
        if (int rc = mx_convert(h, 1)) return rc;  // ts of the state, with the old stage's u
      const double cf_old = h->cf;
      h->cf += cuf;
      if (h->cf > 1.0) h->cf = 1.0;
      const double rem = cuf - (h->cf - cf_old);
      const double dcf = h->cf - cf_old;
      std::vector<double> coef(V), hcoef(V);
      for (int v = 0; v < V; ++v) {
        coef[v] = h->qv[v] * rem;
        hcoef[v] = -h->qv[v] * dcf;  // history keeps the same t in the new frame
      }
      if (rem != 0.0) HIP_OK(launch_axpy_fr<double>(h->ts, h->fr, coef.data(), V, N, 1, h->stream));
      if (!kMixed) {
        HIP_OK(launch_axpy_fr<double>((double *)h->hist_t, h->fr, hcoef.data(), V, N, H, h->stream));
      } else if (c.filled > 0) {
        // increment form: x_(k) - x moves by -q cuf fr for every age k, i.e.
        // the newest increment grows by q cuf fr
        const int s1 = (c.write - 1 + H) % H;
        std::vector<double> inc(V);
        for (int v = 0; v < V; ++v) inc[v] = h->qv[v] * cuf;
        HIP_OK(launch_axpy_fr<float>(hx32 + (size_t)s1 * VN, h->fr, inc.data(), V, N, 1, h->stream));
      }
    }
    if (int rc = stage_potentials(h)) return rc;
    if (kMixed) {  // this stage's reference: A = ts - u, and a refresh first
      if (int rc = mx_convert(h, 0)) return rc;
      h->mx_live = true;
    }
    if (verbose) std::printf("relaxing 3D UV RISM: Charge Up Factor = %g\n", h->cf);
    c.status = -1;
    c.step = 0;
    c.run_iter = c.cur;
    c.upd_iter = -1;
    if (kMixed) c.refresh_iter = c.cur;
    if (int rc = put_ctl(h)) return rc;
    while (true) {
      int n_launch = (h->opts.ksave > 0) ? 1 : chunk;
      if (kMixed && c.refresh_iter == c.cur) {  // the solve kernel asked for a refresh
        if (trace) std::fprintf(stderr, "rism: refresh at step %d (rms %.3e)\n", c.total, c.rms);
        if (int rc = launch_mx_iteration(h, 2)) return rc;
        ++h->refreshes;
        n_launch = 1;
      } else if (graphs && n_launch > 1) {
        // a refresh asked for inside the chunk turns its remaining launches off
        if (int rc = launch_chunk_graph<R>(h, n_launch)) return rc;
      } else {
        for (int k = 0; k < n_launch; ++k)
          if (int rc = launch_iteration<R>(h, false, true)) return rc;
      }
      if (int rc = get_ctl(h)) return rc;
      // progress lines (iterate.cu:63)
      if (c.total > printed) {
        // the control block carries the last kRmsRecent steps' rms (a
        // launch group has at most that many steps); rms_hist the rest
        std::vector<double> tail(c.total - printed);
        if (c.total - printed <= kRmsRecent) {
          for (int k = printed; k < c.total; ++k) tail[k - printed] = c.rms_recent[k % kRmsRecent];
        } else {
          HIP_OK(hcopy(h, tail.data(), h->rms_hist + printed, tail.size() * sizeof(double),
                           hipMemcpyDeviceToHost));
        }
        for (size_t k = 0; k < tail.size(); ++k) {
          h->rms_log.push_back(tail[k]);
          if (verbose)
            std::printf(" Step = %d Reside = %g\n", (int)(c.step - (tail.size() - 1 - k)), tail[k]);
        }
        printed = c.total;
      }
      if (h->opts.ksave > 0 && c.step > 0 && c.step % h->opts.ksave == 0 && c.status != RISM_SOLVE_DIVERGED &&
          tuv_path && tuv_path[0]) {
        // iterate.cu:64-66 writes t after the Anderson update of this step:
        // apply the pending update now (one first pass), then clear it so
        // the next iteration does not apply it again
        if (c.upd_iter == c.cur || (c.status == RISM_SOLVE_MAXSTEP && c.upd_hold == c.cur)) {
          const int keep_run = c.run_iter;
          c.run_iter = -1;  // update only
          c.upd_iter = c.cur;
          if (int rc = put_ctl(h)) return rc;
          if (int rc = launch_update_only(h)) return rc;
          c.upd_iter = -1;
          c.upd_hold = -1;
          c.run_iter = keep_run;
          if (int rc = put_ctl(h)) return rc;
        }
        if (verbose) std::printf("saving Tuv to file:  %s  ...\n", tuv_path);
        if (int rc = write_tuv_file(h, tuv_path)) return rc;
      }
      if (c.status >= 0) break;
    }
    // a step-limit stop still applies its pending update (iterate.cu:61 runs
    // before the loop ends)
    if (c.status == RISM_SOLVE_MAXSTEP && c.upd_hold == c.cur) {
      c.run_iter = -1;  // update only, no closure/FFT
      c.upd_iter = c.cur;
      if (int rc = put_ctl(h)) return rc;
      if (int rc = launch_update_only(h)) return rc;
      c.upd_hold = -1;
    }
    status = c.status;
    if (status == RISM_SOLVE_DIVERGED) {
      if (verbose) std::printf("Calculation diverged.\n");
    } else if (status == RISM_SOLVE_MAXSTEP) {
      if (verbose) std::printf("3D UV RISM: reached limit # of relaxation steps: %d\n", h->opts.maxstep);
      break;
    } else if (h->opts.ksave != 0 && stage == cu && tuv_path && tuv_path[0]) {
      if (verbose) std::printf("saving Tuv to file:  %s  ...\n", tuv_path);
      if (int rc = write_tuv_file(h, tuv_path)) return rc;
    }
  }
  HIP_OK(hipStreamSynchronize(h->stream));
  if (int rc = prof_collect(h)) return rc;
  h->steps = c.total;
  h->rms = c.rms;
  h->solved = true;
  h->t_solve = now_s() - t0;
  if (steps_out) *steps_out = c.total;
  if (rms_out) *rms_out = c.rms;
  return status;
}

// re-evaluate the map at the final state, keeping real-space h in Z and g
template <typename R> int evaluate_impl(rism_handle *h) {
  Ctl &c = *h->ctl_host;
  if (int rc = get_ctl(h)) return rc;
  c.run_iter = c.cur;
  c.upd_iter = -1;
  const int keep = c.status;
  if (int rc = put_ctl(h)) return rc;
  if (int rc = launch_iteration<R>(h, true, false, true)) return rc;
  IterArgs<R, double> a = make_args<R>(h);
  HIP_OK(launch_gfield<R, double>(a, h->gfield, h->stream));
  c.run_iter = -1;
  c.status = keep;
  if (int rc = put_ctl(h)) return rc;
  h->evaluated = true;
  return RISM_OK;
}

// The final state is always re-evaluated with the FP64 passes (the mixed
// path keeps an FP64 field set for its refinement), so h, g and every
// thermodynamic sum come from FP64 fields.
int evaluate(rism_handle *h) {
  if (h->mx_live)  // t_s of the mixed path's final state
    if (int rc = mx_convert(h, 1)) return rc;
  return evaluate_impl<double>(h);
}

template <typename R> int thermo_impl(rism_handle *h, rism_xmu *x) {
  const int V = h->V;
  const int nblk = 1024;
  IterArgs<R, double> a = make_args<R>(h);
  PROF(h, 8, launch_exchem<R, double>(a, h->fr, h->partials, nblk, h->stream));
  std::vector<double> part((size_t)nblk * 4 * V);
  HIP_OK(hipMemcpyAsync(part.data(), h->partials, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  std::vector<double> tot(4 * V, 0.0);
  for (int b = 0; b < nblk; ++b)
    for (int k = 0; k < 4 * V; ++k) tot[k] += part[(size_t)b * 4 * V + k];
  PROF(h, 9, launch_se_sum(h->N, V, h->d_rho, h->d_qv, h->gfield, h->sel, h->seq, h->partials, nblk, h->stream));
  std::vector<double> pe((size_t)nblk * 2);
  HIP_OK(hipMemcpyAsync(pe.data(), h->partials, pe.size() * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  double se_lj = 0.0, se_es = 0.0;
  for (int b = 0; b < nblk; ++b) {
    se_lj += pe[2 * b];
    se_es += pe[2 * b + 1];
  }
  std::memset(x, 0, sizeof *x);
  x->nsite = V;
  const double ibeta = kAvogadro * kBoltzmann * h->temper;
  double sc = 0.0, schnc = 0.0, gf = 0.0, cuv = 0.0;
  for (int v = 0; v < V; ++v) {  // cal_exchem.cu:9-51, output_xmu.cu:31-73
    const double s0 = tot[4 * v + 0] * h->rho[v] * h->dv;
    const double s1 = tot[4 * v + 1] * h->rho[v] * h->dv;
    const double s2 = tot[4 * v + 2] * h->rho[v] * h->dv;
    x->sfec_sc[v] = ibeta * s0;
    x->sfec_sc_hnc[v] = h->opts.closure == RISM_CLOSURE_KH ? ibeta * s1 : 0.0;
    x->sfec_gf[v] = ibeta * s2;
    sc += s0;
    if (h->opts.closure == RISM_CLOSURE_KH) schnc += s1;
    gf += s2;
    cuv += tot[4 * v + 3] * h->rho[v];
  }
  x->sfe_sc = ibeta * sc;
  x->sfe_sc_hnc = ibeta * schnc;
  x->sfe_gf = ibeta * gf;
  x->se_lj = se_lj * h->dv;
  x->se_es = se_es * h->dv;
  x->se = (se_lj + se_es) * h->dv;
  cuv *= h->dv;  // cal_pmv.cu:5-22
  const double pmv = h->xt * (1.0 - cuv) * ibeta;
  double pr = 0.0;  // cal_pressure.cu:6-20
  for (int v = 0; v < V; ++v) pr += h->rho[v];
  pr *= 1.0e30 / kAvogadro;
  pr = 0.5 * (pr + 1.0 / (h->xt * ibeta));
  x->pmv = pmv * 1000.0;
  x->pressure = ibeta * pr;
  x->correction = -pr * pmv * ibeta;
  return RISM_OK;
}

template <typename R> int alloc_precision(rism_handle *h) {
  using T2 = typename C2<R>::T;
  using T4 = typename R4T<R>::T;
  PrecSet &ps = h->ps[pidx<R>()];
  const size_t VN = (size_t)h->V * h->N;
  if (h->tmpl) {  // the first member's plans and tables (not owned)
    const PrecSet &t = h->tmpl->ps[pidx<R>()];
    for (int k = 0; k < 3; ++k) ps.plan[k] = t.plan[k];
    ps.plan_xr = t.plan_xr;
    ps.geom = t.geom;
    ps.tables = t.tables;
    ps.chi = t.chi;
  } else {
    h->atom_alloc = true;  // tables: outside a batch arena, shared by its members
    int rc = build_plans<R>(h);
    h->atom_alloc = false;
    if (rc) return rc;
    if ((rc = choose_geometry(h, ps.geom, sizeof(R)))) return rc;
  }
  // the FP64 set carries the potentials; the mixed path's FP32 chain works
  // on differences to an FP64 reference and needs neither u nor the
  // long-range terms
  if (sizeof(R) == sizeof(double)) {
    if (int rc = dalloc(h, &ps.u, VN * sizeof(R))) return rc;
    ps.fr = h->fr;
    ps.fkh = h->fkh;
  }
  if (int rc = dalloc(h, &ps.Z, (size_t)h->P * h->N * sizeof(T2))) return rc;
  if (h->tmpl) {
    ps.ready = true;
    return RISM_OK;
  }
  // chi spline table [V*V][ntab-1] of (y_i, c0, c1, c2)
  const int nt = h->ntab, V = h->V;
  std::vector<double> kt(nt), y(nt), c0(nt), c1(nt), c2(nt);
  for (int i = 0; i < nt; ++i) kt[i] = h->dk * (i + 1);
  std::vector<T4> tab((size_t)V * V * (nt - 1));
  for (int v = 0; v < V; ++v)
    for (int v1 = 0; v1 < V; ++v1) {
      for (int i = 0; i < nt; ++i) y[i] = h->chi[((size_t)i * V + v) * V + v1];
      natural_spline(kt.data(), y.data(), nt, c0.data(), c1.data(), c2.data());
      for (int i = 0; i < nt - 1; ++i) {
        T4 &e = tab[(size_t)(v * V + v1) * (nt - 1) + i];
        e.x = (R)y[i];
        e.y = (R)c0[i];
        e.z = (R)c1[i];
        e.w = (R)c2[i];
      }
    }
  h->atom_alloc = true;
  const int rcc = dalloc(h, &ps.chi, tab.size() * sizeof(T4));
  h->atom_alloc = false;
  if (rcc) return rcc;
  HIP_OK(hcopy(h, ps.chi, tab.data(), tab.size() * sizeof(T4), hipMemcpyHostToDevice));
  ps.ready = true;
  return RISM_OK;
}

int upload_solute(rism_handle *h);

int create_impl(const rism_grid *grid, const rism_solvent *sv, const rism_solute *su, const rism_opts *o,
                int device, rism_handle *h) {
  const double t0 = now_s();
  for (int k = 0; k < 3; ++k) {
    if (!is_supported_length(grid->n[k])) {
      set_error("grid length " + std::to_string(grid->n[k]) +
                " unsupported: need even, 2,3,5-smooth, 4..1024");
      return RISM_ERR_ARG;
    }
    if (!(grid->box[k] > 0.0)) {
      set_error("box lengths must be positive");
      return RISM_ERR_ARG;
    }
  }
  if (sv->nsite < 1 || sv->nsite > kMaxV || sv->ntab < 4 || !(sv->dk > 0.0)) {
    set_error("bad solvent description (1..8 sites, ntab >= 4, dk > 0)");
    return RISM_ERR_ARG;
  }
  if (o->nhist < 1 || o->nhist > 10) {
    set_error("nhist must be 1..10");
    return RISM_ERR_ARG;
  }
  if (o->closure != RISM_CLOSURE_KH && o->closure != RISM_CLOSURE_HNC) {
    set_error("unknown closure");
    return RISM_ERR_ARG;
  }
  if (o->precision != RISM_PREC_FP64 && o->precision != RISM_PREC_MIXED) {
    set_error("unknown precision");
    return RISM_ERR_ARG;
  }
  if (o->maxstep < 1) {
    set_error("maxstep must be >= 1");
    return RISM_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no HIP device available");
    return RISM_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    set_error("device index out of range");
    return RISM_ERR_ARG;
  }
  h->device = device;
  HIP_OK(hipSetDevice(device));
  {
    std::lock_guard<std::mutex> lk(g_device_mu);
    for (size_t i = 0; i < g_spare.size() && !h->borrowed_ctx; ++i)
      if (g_spare[i].device == device) {
        h->stream = g_spare[i].stream;
        h->ctl_host = g_spare[i].ctl_host;
        g_spare[i] = g_spare.back();
        g_spare.pop_back();
        break;
      }
    if (!h->stream) HIP_OK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    if (!h->ctl_host) HIP_OK(hipHostMalloc((void **)&h->ctl_host, sizeof(Ctl), hipHostMallocDefault));
  }
  h->nx = grid->n[0]; h->ny = grid->n[1]; h->nz = grid->n[2];
  h->N = (long)h->nx * h->ny * h->nz;
  for (int k = 0; k < 3; ++k) {  // cell.cc:8-22
    h->box[k] = grid->box[k];
    h->dr[k] = grid->box[k] / grid->n[k];
  }
  h->volume = h->box[0] * h->box[1] * h->box[2];
  h->dv = h->volume / (double)h->N;
  h->V = sv->nsite;
  h->P = (h->V + 1) / 2;
  h->rho.assign(sv->rho, sv->rho + h->V);
  h->qv.assign(sv->q, sv->q + h->V);
  h->sigv.assign(sv->sig, sv->sig + h->V);
  h->epsv.assign(sv->eps, sv->eps + h->V);
  h->temper = sv->temper;
  h->xt = sv->xt;
  h->dk = sv->dk;
  h->ntab = sv->ntab;
  h->chi.assign(sv->chi, sv->chi + (size_t)sv->ntab * h->V * h->V);
  h->natu = su->natom;
  h->usig.assign(su->sig, su->sig + su->natom);
  h->ueps.assign(su->eps, su->eps + su->natom);
  h->uq.assign(su->q, su->q + su->natom);
  h->uxyz.assign(su->xyz, su->xyz + 3 * (size_t)su->natom);
  h->opts = *o;
  h->hm = o->nhist <= 2 ? 2 : (o->nhist <= 5 ? 5 : 10);
  h->qsz = sizeof(double);  // residual / history width of the FP64 path
  // k range check (solvent_spline.cu:18-21)
  double kmax2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double km = 2.0 * M_PI * (grid->n[k] - 1 - grid->n[k] / 2.0 + 0.5) / grid->box[k];
    kmax2 += km * km;
  }
  if (std::sqrt(kmax2) > h->dk * h->ntab) {
    set_error("insufficient maximal T tabulated");
    return RISM_ERR_RANGE;
  }
  const size_t VN = (size_t)h->V * h->N;
  const size_t nh = (size_t)h->nx * (h->ny / 2) * h->nz;
  int rc = 0;
  if ((rc = dalloc_t(h, &h->ulj, VN)) || (rc = dalloc_t(h, &h->sel, 2 * VN)) ||
      (rc = dalloc_t(h, &h->de, h->N)) || (rc = dalloc_t(h, &h->fr, h->N)) ||
      (rc = dalloc_t(h, &h->sr, h->N)) || (rc = dalloc_t(h, &h->fkh, nh)) ||
      (rc = dalloc_t(h, &h->ts, VN)) || (rc = dalloc_t(h, &h->gfield, VN)) || (rc = dalloc_t(h, &h->ctl, 1)) ||
      (rc = dalloc_t(h, &h->rms_hist, (size_t)(o->maxstep + 1) * (o->chargeup + 1) + 4)) ||
      (rc = dalloc_t(h, &h->d_qv, h->V)) || (rc = dalloc_t(h, &h->d_rho, h->V)))
    return rc;
  h->atom_alloc = true;  // solute-sized: never from a batch arena
  rc = dalloc_t(h, &h->phase, (size_t)std::max(1, h->natu) * (h->nx + h->ny + h->nz));
  if (!rc) rc = dalloc_t(h, &h->atoms, std::max(1, h->natu));
  if (!rc) rc = dalloc_t(h, &h->lj, (size_t)h->V * std::max(1, h->natu));
  if (!rc) rc = dalloc_t(h, &h->sigh, (size_t)h->V * std::max(1, h->natu));
  h->atom_alloc = false;
  if (rc) return rc;
  h->seq = h->sel + VN;  // the two SE atom sums share one allocation
  std::memset(h->ctl_host, 0, sizeof(Ctl));
  if ((rc = alloc_precision<double>(h))) return rc;
  if (o->precision == RISM_PREC_MIXED) {
    // reference T (FP64), A, change, R_ref, B (FP32), increment history and
    // the residual ring of H + 1 slots (FP32)
    if ((rc = alloc_precision<float>(h)) || (rc = dalloc_t(h, &h->mT, VN)) || (rc = dalloc_t(h, &h->mA, VN)) ||
        (rc = dalloc_t(h, &h->mdl, VN)) ||
        (rc = dalloc_t(h, &h->mrref, VN)) || (rc = dalloc_t(h, &h->mbq, VN)) ||
        (rc = dalloc_t(h, &h->mhx, (size_t)o->nhist * VN)) || (rc = dalloc_t(h, &h->mring, (size_t)(o->nhist + 1) * VN)))
      return rc;
  } else if ((rc = dalloc(h, &h->res, VN * h->qsz)) || (rc = dalloc(h, &h->hist_t, (size_t)o->nhist * VN * h->qsz)) ||
             (rc = dalloc(h, &h->hist_r, (size_t)o->nhist * VN * h->qsz))) {
    return rc;
  }
  {  // per-workgroup partial sums: residual pass and thermodynamic reductions
    const size_t nv = 1 + h->hm + h->hm * (h->hm + 1) / 2;
    const size_t npart = std::max<size_t>((size_t)1024 * 4 * kMaxV, (size_t)kMaxXBlocks * nv);
    if ((rc = dalloc_t(h, &h->partials, npart))) return rc;
  }
  h->natu_cap = std::max(1, h->natu);
  if ((rc = upload_solute(h))) return rc;
  h->t_setup = now_s() - t0;
  return RISM_OK;
}

// solute records and Lorentz-Berthelot pair parameters (cal_LJ.cu:24-31)
int upload_solute(rism_handle *h) {
  std::vector<double4> at(std::max(1, h->natu));
  std::vector<double2> lj((size_t)h->V * std::max(1, h->natu));
  std::vector<double> sigh((size_t)h->V * std::max(1, h->natu));
  for (int u = 0; u < h->natu; ++u) at[u] = make_double4(h->uxyz[3 * u], h->uxyz[3 * u + 1], h->uxyz[3 * u + 2], h->uq[u]);
  for (int v = 0; v < h->V; ++v)
    for (int u = 0; u < h->natu; ++u) {
      const double s = (h->usig[u] + h->sigv[v]) * 0.5;
      const double e = std::sqrt(h->ueps[u] * h->epsv[v]);
      lj[(size_t)v * h->natu + u] = make_double2(s * s, e * 4.0);
      sigh[(size_t)v * h->natu + u] = s * 0.5;
    }
  HIP_OK(hcopy(h, h->atoms, at.data(), at.size() * sizeof(double4), hipMemcpyHostToDevice));
  HIP_OK(hcopy(h, h->lj, lj.data(), lj.size() * sizeof(double2), hipMemcpyHostToDevice));
  HIP_OK(hcopy(h, h->sigh, sigh.data(), sigh.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hcopy(h, h->d_qv, h->qv.data(), h->V * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hcopy(h, h->d_rho, h->rho.data(), h->V * sizeof(double), hipMemcpyHostToDevice));
  return RISM_OK;
}

}  // namespace

// A new solute on an existing handle (same grid, solvent and options): the
// field buffers and the captured iteration graphs stay; the solute-sized
// arrays grow when needed (the adjoint's are then allocated afresh on its
// next use).  Potentials must be built again before the next solve.
int rism::set_solute(rism_handle *h, const rism_solute *su) {
  if (!h || !su || su->natom < 0 || (su->natom > 0 && (!su->sig || !su->eps || !su->q || !su->xyz))) {
    set_error("set_solute: bad arguments");
    return RISM_ERR_ARG;
  }
  HIP_OK(hipSetDevice(h->device));
  const int nu = su->natom;
  if (std::max(1, nu) > h->natu_cap) {
    const int cap = std::max(1, nu);
    // the superseded arrays go back to the cache
    dfree(h, h->atoms);
    dfree(h, h->lj);
    dfree(h, h->sigh);
    dfree(h, h->phase);
    h->atoms = nullptr;
    h->lj = nullptr;
    h->sigh = nullptr;
    h->phase = nullptr;
    h->natu_cap = 0;
    int rc = 0;
    h->atom_alloc = true;
    if ((rc = dalloc_t(h, &h->atoms, cap)) || (rc = dalloc_t(h, &h->lj, (size_t)h->V * cap)) ||
        (rc = dalloc_t(h, &h->sigh, (size_t)h->V * cap)) ||
        (rc = dalloc_t(h, &h->phase, (size_t)cap * (h->nx + h->ny + h->nz)))) {
      h->atom_alloc = false;
      return rc;
    }
    h->atom_alloc = false;
    h->natu_cap = cap;
  }
  h->natu = nu;
  h->usig.assign(su->sig, su->sig + nu);
  h->ueps.assign(su->eps, su->eps + nu);
  h->uq.assign(su->q, su->q + nu);
  h->uxyz.assign(su->xyz, su->xyz + 3 * (size_t)nu);
  if (int rc = upload_solute(h)) return rc;
  h->potentials = h->solved = h->evaluated = false;
  return RISM_OK;
}

// ---------------------------------------------------------------------
// Batched solves: many solutes of one grid class (same grid, box, solvent
// and options) in one launch per kernel.  Every member handle carves its
// grid-sized buffers from its slice of one arena, in the same order, so
// each field of solute s lies s * stride bytes after solute 0's and one
// launch with blockIdx.y = s covers them all (IterArgs::astride).  Each
// solute keeps its own control block, which gates its workgroups: that is
// its convergence mask.  A chunk of iterations (in the mixed path both
// the FP32 chain and the FP64 refresh chain, each running only for the
// solutes whose step it is) is captured once as a HIP graph and replayed
// until every solute has stopped.
namespace {

// One step of every solute of a batch.  Mixed precision: the FP64 refresh
// chain, then the FP32 chain.  With a refresh list (rlist, capacity cap)
// the refresh chain launches cap rows of workgroups for the solutes on the
// list, and the FP32 chain's solve files the next step's refreshes there;
// at the first step (all_refresh: every solute refreshes) or without a list
// it launches all S rows, each gated by its solute's control block.
int batch_iteration(rism_handle *h0, int S, long stride, int *rlist = nullptr, int cap = 0,
                    bool all_refresh = false) {
  const bool mixed = h0->opts.precision == RISM_PREC_MIXED;
  if (mixed) {
    {  // FP64 refresh chain: the solutes whose refresh step this is
      IterArgs<double, double> a = make_args<double>(h0, true);
      const LaunchGeom &g = h0->ps[0].geom;
      const bool listed = rlist && !all_refresh;
      const int rows = listed ? cap : S;
      a.gate = 2;
      a.use_fk = 1;
      a.astride = stride;
      a.nsol = rows;
      a.rlist = listed ? rlist : nullptr;
      HIP_OK(launch_mx_closure_fx<double>(a, g, h0->hm, h0->stream));
      HIP_OK(launch_fft_y<double, double>(a, g, false, h0->stream));
      HIP_OK(launch_fft_z_oz<double, double>(a, g, h0->stream));
      HIP_OK(launch_fft_y<double, double>(a, g, true, h0->stream));
      HIP_OK(launch_mx_resid<double>(a, g, h0->hm, h0->stream));
      HIP_OK(launch_anderson_solve(h0->ctl, h0->partials, g.xblocks, a.nv, h0->hm, h0->rms_hist, h0->stream, 2,
                                   stride, rows, a.rlist));
    }
    IterArgs<float, double> a = make_args<float>(h0);
    const LaunchGeom &g = h0->ps[1].geom;
    a.gate = 1;
    a.use_fk = 0;
    a.astride = stride;
    a.nsol = S;
    a.rlist = rlist;  // emptied by the closure pass, filled by the solve
    HIP_OK(launch_mx_closure_fx<float>(a, g, h0->hm, h0->stream));
    HIP_OK(launch_fft_y<float, double>(a, g, false, h0->stream));
    HIP_OK(launch_fft_z_oz<float, double>(a, g, h0->stream));
    HIP_OK(launch_fft_y<float, double>(a, g, true, h0->stream));
    HIP_OK(launch_mx_resid<float>(a, g, h0->hm, h0->stream));
    HIP_OK(launch_anderson_solve(h0->ctl, h0->partials, g.xblocks, a.nv, h0->hm, h0->rms_hist, h0->stream, 1,
                                 stride, S, rlist));
    return RISM_OK;
  }
  IterArgs<double, double> a = make_args<double>(h0);
  const LaunchGeom &g = h0->ps[0].geom;
  a.astride = stride;
  a.nsol = S;
  HIP_OK(launch_closure_fx<double, double>(a, g, true, h0->hm, h0->stream));
  HIP_OK(launch_fft_y<double, double>(a, g, false, h0->stream));
  HIP_OK(launch_fft_z_oz<double, double>(a, g, h0->stream));
  HIP_OK(launch_fft_y<double, double>(a, g, true, h0->stream));
  HIP_OK(launch_fx_resid<double, double>(a, g, h0->hm, h0->stream));
  HIP_OK(launch_anderson_solve(h0->ctl, h0->partials, g.xblocks, a.nv, h0->hm, h0->rms_hist, h0->stream, 0, stride,
                               S));
  return RISM_OK;
}

// the per-solute start of a solve (solve_impl's set-up, cu = 0, no restart
// file) on the batch's stream
int batch_member_start(rism_handle *h, hipStream_t st) {
  const int V = h->V, H = h->opts.nhist;
  const size_t VN = (size_t)V * h->N;
  const bool mixed = h->opts.precision == RISM_PREC_MIXED;
  hipStream_t own = h->stream;
  h->stream = st;  // every launch of the batch is ordered on one stream
  h->evaluated = false;
  h->mx_live = false;
  h->cf = 1.0;
  if (mixed) {
    HIP_OK(hipMemsetAsync(h->mhx, 0, (size_t)H * VN * sizeof(float), st));
    HIP_OK(hipMemsetAsync(h->mring, 0, (size_t)(H + 1) * VN * sizeof(float), st));
  } else {
    HIP_OK(hipMemsetAsync(h->hist_t, 0, (size_t)H * VN * h->qsz, st));
    HIP_OK(hipMemsetAsync(h->hist_r, 0, (size_t)H * VN * h->qsz, st));
    HIP_OK(hipMemsetAsync(h->res, 0, VN * h->qsz, st));
  }
  HIP_OK(hipMemsetAsync(h->ts, 0, VN * sizeof(double), st));  // initialize_tuv.cu: t = q_v fr
  int rc = stage_potentials(h);
  if (!rc && mixed) {
    rc = mx_convert(h, 0);
    h->mx_live = true;
  }
  Ctl &c = *h->ctl_host;
  std::memset(&c, 0, sizeof c);
  c.count = h->opts.count;
  c.nhist = H;
  c.maxstep = h->opts.maxstep;
  c.tol = h->opts.tol;
  c.m = h->opts.m;
  c.mp = h->opts.mp;
  c.npts = (double)VN;
  c.status = -1;
  c.run_iter = 0;
  c.upd_iter = -1;
  c.refresh_iter = mixed ? 0 : -1;
  if (mixed) {
    c.ring = 1;
    c.rcur = H;
    for (int k = 0; k < H; ++k) c.rmap[k] = k;
  }
  if (!rc && hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice, st) != hipSuccess) {
    set_error("batch: control block upload failed");
    rc = RISM_ERR_HIP;
  }
  h->stream = own;
  return rc;
}

}  // namespace

// S solutes on one grid: create, potentials, one batched solve, thermo.
int rism::batch_solve_class(const rism_grid *grid, const rism_solvent *sv, const rism_solute *const *sus, int S,
                            const rism_opts *o, int device, int *status, int *steps, double *rms, rism_xmu *xmu) {
  for (int s = 0; s < S; ++s) status[s] = RISM_ERR_STATE;
  if (S <= 0) return RISM_OK;
  if (o->chargeup != 0 || o->ksave != 0) {
    set_error("batched solves need chargeup = 0 and ksave = 0");
    for (int s = 0; s < S; ++s) status[s] = RISM_ERR_ARG;
    return RISM_ERR_ARG;
  }
  const bool trace = env_long("RISM_TRACE", 0) > 0;
  const double tb0 = now_s();
  // 1. the slice size: one member created normally, its grid-sized buffers counted
  size_t stride = 0;
  {
    rism_handle *m = new rism_handle();
    m->arena_measure = true;
    int rc = create_impl(grid, sv, sus[0], o, device, m);
    stride = m->arena_need;
    const std::string err = rc ? std::string(rism::last_error()) : std::string();
    rism_destroy(m);
    if (rc) {
      set_error(err);
      for (int s = 0; s < S; ++s) status[s] = rc;
      return rc;
    }
  }
  HIP_OK(hipSetDevice(device));
  char *arena = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_device_mu);
    hipError_t e = hipMalloc((void **)&arena, stride * (size_t)S);
    if (e != hipSuccess && !g_cache.empty()) {
      (void)hipGetLastError();
      cache_flush_locked(device);
      e = hipMalloc((void **)&arena, stride * (size_t)S);
    }
    if (e != hipSuccess) {
      set_error("batch arena of " + std::to_string(stride * (size_t)S) + " bytes: " + hipGetErrorString(e));
      for (int s = 0; s < S; ++s) status[s] = RISM_ERR_NOMEM;
      return RISM_ERR_NOMEM;
    }
  }
  std::vector<rism_handle *> hs(S, nullptr);
  int rc = RISM_OK;
  // one pinned block holds the members' control blocks (the first member's is its own)
  Ctl *pinned = nullptr;
  if (S > 1 && hipHostMalloc((void **)&pinned, sizeof(Ctl) * (size_t)S, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    pinned = nullptr;
  }
  for (int s = 0; s < S && rc == RISM_OK; ++s) {
    rism_handle *h = new rism_handle();
    h->arena = arena + (size_t)s * stride;
    h->arena_cap = stride;
    h->tmpl = s > 0 ? hs[0] : nullptr;
    if (s > 0 && pinned) {  // the batch's stream and a slot of the pinned block
      h->stream = hs[0]->stream;
      h->ctl_host = pinned + s;
      h->borrowed_ctx = true;
    }
    rc = create_impl(grid, sv, sus[s], o, device, h);
    if (rc == RISM_OK && h->arena_used != stride) {
      set_error("batch member's buffers differ from the measured slice");
      rc = RISM_ERR_STATE;
    }
    h->arena = nullptr;  // later buffers (adjoint, scratch) are the handle's own
    hs[s] = h;
  }
  rism_handle *h0 = hs[0];
  hipStream_t st = h0 ? h0->stream : nullptr;
  const bool mixed = o->precision == RISM_PREC_MIXED;
  const double tb1 = now_s();
  // 2. potentials, then every member's start of the solve on the batch's stream
  for (int s = 0; s < S && rc == RISM_OK; ++s) rc = rism_build_potentials(hs[s]);
  const double tb2 = now_s();
  if (rc == RISM_OK && env_long("RISM_AUTOTUNE", 0) > 0 && !h0->tuned) rc = rism_autotune(h0, (int)env_long("RISM_AUTOTUNE", 0), nullptr);
  if (rc == RISM_OK)  // every member launches with solute 0's tile shapes
    for (int s = 1; s < S; ++s) {
      hs[s]->ps[0].geom = h0->ps[0].geom;
      if (mixed) hs[s]->ps[1].geom = h0->ps[1].geom;
    }
  for (int s = 0; s < S && rc == RISM_OK; ++s) rc = batch_member_start(hs[s], st);
  if (rc == RISM_OK && hipStreamSynchronize(st) != hipSuccess) {
    set_error("batch: start of the solves failed");
    rc = RISM_ERR_HIP;
  }
  // 3. the batched solve: chunks of iterations as one graph, replayed until every solute stopped
  const double t0 = now_s();
  const int chunk = 16;
  hipGraphExec_t ex = nullptr;
  // mixed: the refresh list, {count, capacity, solute ...}; a solute
  // refreshes about once in 30 steps, so capacity S/8 (RISM_BATCH_REFRESH_DIV)
  // leaves room for bursts, and a solute that finds the list full refreshes
  // a step later
  int *rlist = nullptr;
  int rcap = 0;
  if (rc == RISM_OK && mixed && S > 1) {
    const long div = std::max(1L, env_long("RISM_BATCH_REFRESH_DIV", 8));
    rcap = (int)std::max(1L, ((long)S + div - 1) / div);
    std::vector<int> init(S + 2, 0);
    init[1] = rcap;
    {
      std::lock_guard<std::mutex> lk(g_device_mu);
      if (hipMalloc((void **)&rlist, sizeof(int) * (S + 2)) != hipSuccess) {
        (void)hipGetLastError();
        rlist = nullptr;  // no list: every refresh launch covers all S solutes
      }
    }
    if (rlist && hipMemcpy(rlist, init.data(), sizeof(int) * (S + 2), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("batch: refresh list upload failed");
      rc = RISM_ERR_HIP;
    }
    // the first step is a refresh for every solute: all S rows
    if (rc == RISM_OK) rc = batch_iteration(h0, S, (long)stride, rlist, rcap, true);
  }
  if (rc == RISM_OK) {
    const char *genv = std::getenv("RISM_GRAPHS");
    bool graphs = !(genv && genv[0] == '0');
    if (graphs) {
      std::unique_lock<std::mutex> lk(g_device_mu);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
      int r2 = RISM_OK;
      if (e == hipSuccess) {
        for (int k = 0; k < chunk && r2 == RISM_OK; ++k) r2 = batch_iteration(h0, S, (long)stride, rlist, rcap);
        e = hipStreamEndCapture(st, &g);
      }
      if (e == hipSuccess && r2 == RISM_OK) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (e != hipSuccess || r2 != RISM_OK) {
        (void)hipGetLastError();
        if (ex) (void)hipGraphExecDestroy(ex);
        ex = nullptr;
      }
    }
    int live = S;
    long guard = 0;
    while (live > 0 && rc == RISM_OK) {
      if (ex) {
        if (hipGraphLaunch(ex, st) != hipSuccess) { set_error("batch: graph launch failed"); rc = RISM_ERR_HIP; break; }
      } else {
        for (int k = 0; k < chunk && rc == RISM_OK; ++k) rc = batch_iteration(h0, S, (long)stride, rlist, rcap);
      }
      for (int s = 0; s < S && rc == RISM_OK; ++s)
        if (hipMemcpyAsync(hs[s]->ctl_host, hs[s]->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st) != hipSuccess) {
          set_error("batch: control block read failed");
          rc = RISM_ERR_HIP;
        }
      if (rc == RISM_OK && hipStreamSynchronize(st) != hipSuccess) { set_error("batch: solve failed"); rc = RISM_ERR_HIP; }
      live = 0;
      for (int s = 0; s < S; ++s) live += hs[s]->ctl_host->status < 0;
      if (++guard > (long)o->maxstep / chunk + 4) break;  // every solute stops within maxstep steps
    }
  }
  if (ex) (void)hipGraphExecDestroy(ex);
  if (rlist) {
    (void)hipStreamSynchronize(st);
    std::lock_guard<std::mutex> lk(g_device_mu);
    (void)hipFree(rlist);
  }
  const double t_solve = now_s() - t0;
  const double tb3 = now_s();
  // 4. per solute: a step-limit stop applies its pending update, then thermo
  for (int s = 0; s < S; ++s) {
    rism_handle *h = hs[s];
    if (!h || rc != RISM_OK) { status[s] = rc ? rc : RISM_ERR_STATE; continue; }
    Ctl &c = *h->ctl_host;
    int r = RISM_OK;
    if (c.status == RISM_SOLVE_MAXSTEP && c.upd_hold == c.cur) {
      c.run_iter = -1;
      c.upd_iter = c.cur;
      r = put_ctl(h);
      if (!r) r = launch_update_only(h);
    }
    h->steps = c.total;
    h->rms = c.rms;
    h->solved = true;
    h->t_solve = t_solve;
    if (steps) steps[s] = c.total;
    if (rms) rms[s] = c.rms;
    const int st_s = c.status < 0 ? RISM_ERR_STATE : c.status;
    if (!r && xmu && st_s >= 0) r = rism_thermo(h, &xmu[s]);
    status[s] = r < 0 ? r : st_s;
  }
  const double tb4 = now_s();
  for (int s = S - 1; s >= 0; --s)  // the first member, whose tables and stream the others share, last
    if (hs[s]) rism_destroy(hs[s]);
  if (pinned) (void)hipHostFree(pinned);
  {
    std::lock_guard<std::mutex> lk(g_device_mu);
    (void)hipFree(arena);
  }
  if (trace)
    std::fprintf(stderr, "rism batch %d x %dx%dx%d: create %.3f s, potentials %.3f, start %.3f, solve %.3f, thermo %.3f, "
                 "release %.3f\n", S, grid->n[0], grid->n[1], grid->n[2], tb1 - tb0, tb2 - tb1, t0 - tb2, tb3 - t0,
                 tb4 - tb3, now_s() - tb4);
  return rc;
}

// ---------------------------------------------------------------------
// Static device scratch of rism_grad_noalloc: one copy per GPU, held until
// the process ends.  In host code `scratch` is a stand-in for the symbol,
// not a device address; each device's own address comes from
// hipGetSymbolAddress, called once per device with that device current,
// and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[RISM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

int grad_scratch_ptr(int device, void **p) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device index out of range for the static scratch");
    return RISM_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {  // the caller has made `device` current
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("hipGetSymbolAddress of the static scratch failed: ") + hipGetErrorString(e));
      return RISM_ERR_HIP;
    }
    g_scratch_ptr[device] = ptr;
  }
  *p = g_scratch_ptr[device];
  return RISM_OK;
}
}  // namespace

// =====================================================================
// C ABI
// =====================================================================
extern "C" {

const char *rism_last_error(void) { return rism::last_error(); }
const char *rism_version(void) { return "rism3d_amd 0.1 (gfx950)"; }

int rism_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int rism_create(const rism_grid *grid, const rism_solvent *solvent, const rism_solute *solute,
                const rism_opts *opts, int device, rism_handle **out) {
  if (!grid || !solvent || !solute || !opts || !out) {
    set_error("null argument");
    return RISM_ERR_ARG;
  }
  *out = nullptr;
  rism_handle *h = new rism_handle();
  int rc = create_impl(grid, solvent, solute, opts, device, h);
  if (rc != RISM_OK) {
    std::string msg = rism::last_error();
    rism_destroy(h);
    set_error(msg);
    return rc;
  }
  *out = h;
  return RISM_OK;
}

int rism_build_potentials(rism_handle *h) {
  if (!h) { set_error("null handle"); return RISM_ERR_ARG; }
  const double t0 = now_s();
  HIP_OK(hipSetDevice(h->device));
  PotArgs a;
  std::memset(&a, 0, sizeof a);
  a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.N = h->N;
  a.V = h->V; a.natu = h->natu;
  for (int k = 0; k < 3; ++k) { a.dr[k] = h->dr[k]; a.box[k] = h->box[k]; }
  a.atoms = h->atoms;
  a.lj = h->lj;
  a.sigh = h->sigh;
  a.sel = h->sel;
  a.seq = h->seq;
  a.ikbt = 1.0 / (kAvogadro * kBoltzmann * h->temper);
  a.ubeta = kHartree * kBohr / (kBoltzmann * h->temper);
  a.ulj = h->ulj; a.de = h->de; a.fr = h->fr; a.sr = h->sr;
  PROF(h, 6, launch_potentials(a, h->stream));
  PROF(h, 7, launch_fk_half(h->nx, h->ny, h->nz, h->natu, h->atoms, h->box, a.ubeta, h->phase, h->fkh, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  if (int rc = prof_collect(h)) return rc;
  h->potentials = true;
  h->solved = h->evaluated = false;
  h->t_setup += now_s() - t0;
  return RISM_OK;
}

int rism_set_esp(rism_handle *h, const double *esp) {
  // cal_Coulomb.cu:40-96: the electrostatic potential of the solute on the
  // grid (Hartree/e) replaces the point-charge sum in de; fr and fk stay the
  // point-charge long-range parts, so the closure exponent's de - fr changes.
  if (!h || !esp) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->potentials) { set_error("rism_build_potentials must run before rism_set_esp"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  const double ub = kHartree / (kBoltzmann * h->temper);
  std::vector<double> de((size_t)h->N), fr((size_t)h->N), sr((size_t)h->N);
  HIP_OK(hcopy(h, fr.data(), h->fr, fr.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (long ip = 0; ip < h->N; ++ip) {
    de[ip] = esp[ip] * ub;
    sr[ip] = de[ip] - fr[ip];
  }
  HIP_OK(hcopy(h, h->de, de.data(), de.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hcopy(h, h->sr, sr.data(), sr.size() * sizeof(double), hipMemcpyHostToDevice));
  h->solved = h->evaluated = false;
  return RISM_OK;
}

}  // extern "C"

// ---- tile autotuner -----------------------------------------------------
// Times one iteration (the five field passes and the solve kernel, without
// the Anderson update, on the initial estimate t = q_v fr) for candidate
// tile shapes and keeps the fastest: the residual pass's x-line group and
// workgroup count, then the closure pass's x-line group, the y tile and the
// z tile, each around the current best.  Results are cached per grid shape, site
// count, history class and precision for the life of the process, so
// batches of solutes on a few grids tune once per grid.
namespace {
struct TuneKey {
  int nx, ny, nz, V, hm, prec;
  bool operator<(const TuneKey &o) const {
    return std::tie(nx, ny, nz, V, hm, prec) < std::tie(o.nx, o.ny, o.nz, o.V, o.hm, o.prec);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::array<int, 5>> g_tune_cache;

template <typename R> int time_iteration(rism_handle *h, int reps, double *ms) {
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  double tot = 0.0;
  int rc = RISM_OK;
  for (int r = -1; r < reps && rc == RISM_OK; ++r) {  // r = -1: warm-up
    Ctl &c = *h->ctl_host;
    std::memset(&c, 0, sizeof c);
    c.nhist = h->opts.nhist;
    c.maxstep = 1;
    c.npts = (double)h->V * h->N;
    c.status = -1;
    c.upd_iter = -1;
    c.refresh_iter = -1;
    if (h->opts.precision == RISM_PREC_MIXED) {  // the FP32 chain's step
      c.ring = 1;
      c.rcur = h->opts.nhist;
      for (int k = 0; k < h->opts.nhist; ++k) c.rmap[k] = k;
    }
    if ((rc = put_ctl(h))) break;
    if (hipEventRecord(e0, h->stream) != hipSuccess) { rc = RISM_ERR_HIP; break; }
    if ((rc = launch_iteration<R>(h, false, true))) break;
    if (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = RISM_ERR_HIP; break; }
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) { rc = RISM_ERR_HIP; break; }
    if (r >= 0) tot += t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == RISM_ERR_HIP) set_error("HIP error while timing a tile candidate");
  *ms = tot / std::max(1, reps);
  return rc;
}

template <typename R> int autotune_impl(rism_handle *h, int reps) {
  PrecSet &ps = h->ps[pidx<R>()];
  const TuneKey key{h->nx, h->ny, h->nz, h->V, h->hm, (int)sizeof(R)};
  std::array<int, 5> best{ps.geom.Lx, ps.geom.Ty, ps.geom.Tz, ps.geom.xblocks, ps.geom.Lc};
  // one tuning at a time: handles on other threads that want the same shape
  // wait here and then take the cached result (no allocation happens under
  // this lock)
  std::lock_guard<std::mutex> lk(g_tune_mu);
  {
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      LaunchGeom g;
      if (choose_geometry(h, g, sizeof(R), it->second.data()) == RISM_OK) ps.geom = g;
      drop_graphs(h);
      return RISM_OK;
    }
  }
  HIP_OK(hipMemsetAsync(h->ts, 0, (size_t)h->V * h->N * sizeof(double), h->stream));
  if (h->opts.precision == RISM_PREC_MIXED) {  // a finite state for the FP32 chain to time
    const size_t VN = (size_t)h->V * h->N;
    HIP_OK(hipMemsetAsync(h->mA, 0, VN * sizeof(float), h->stream));
    HIP_OK(hipMemsetAsync(h->mdl, 0, VN * sizeof(float), h->stream));
    HIP_OK(hipMemsetAsync(h->mrref, 0, VN * sizeof(float), h->stream));
    h->mx_live = false;
  }
  const double cf_save = h->cf;
  h->cf = 1.0;
  auto trial = [&](std::array<int, 5> c, double *ms) -> int {
    LaunchGeom g;
    if (choose_geometry(h, g, sizeof(R), c.data()) != RISM_OK) return 1;         // does not fit: skip
    if (g.Lx != c[0] || g.Ty != c[1] || g.Tz != c[2] || g.Lc != c[4]) return 1;  // not a divisor: skip
    const LaunchGeom keep = ps.geom;
    ps.geom = g;
    int rc = time_iteration<R>(h, reps, ms);
    ps.geom = keep;
    return rc < 0 ? rc : 0;
  };
  double tbest = 0.0;
  auto consider = [&](std::array<int, 5> c) -> int {
    double t = 0.0;
    const int rc = trial(c, &t);
    if (rc < 0) return rc;
    if (rc == 0 && t < tbest) {
      tbest = t;
      best = c;
    }
    return RISM_OK;
  };
  // one factor at a time around the best so far: {index, factors}
  auto around = [&](int idx) -> int {
    for (int f : {2, 4})
      for (int dir : {-1, 1}) {
        std::array<int, 5> c = best;
        c[idx] = dir < 0 ? best[idx] / f : best[idx] * f;
        if (c[idx] >= 1)
          if (int rc = consider(c)) return rc;
      }
    return RISM_OK;
  };
  auto search = [&]() -> int {
    if (int rc = stage_potentials(h)) return rc;
    if (int rc = trial(best, &tbest)) return rc < 0 ? rc : RISM_OK;
    const std::array<int, 5> base = best;
    for (int lf : {1, 2, 4})  // residual-pass x-line group and workgroup count
      for (int xb : {512, 1024, 2048}) {
        if (lf == 1 && xb == base[3]) continue;
        for (int dir : {-1, 1}) {
          if (lf == 1 && dir > 0) continue;
          std::array<int, 5> c = base;
          c[0] = dir < 0 ? base[0] / lf : base[0] * lf;
          c[3] = xb;
          if (c[0] >= 1)
            if (int rc = consider(c)) return rc;
        }
      }
    for (int idx : {4, 1, 2})  // closure-pass x-line group, y tile, z tile
      if (int rc = around(idx)) return rc;
    return RISM_OK;
  };
  const int rc = search();
  h->cf = cf_save;
  if (rc < 0) return rc;
  LaunchGeom g;
  if (int rc2 = choose_geometry(h, g, sizeof(R), best.data())) return rc2;
  ps.geom = g;
  drop_graphs(h);
  h->solved = h->evaluated = false;
  g_tune_cache[key] = best;
  return RISM_OK;
}
}  // namespace

extern "C" {

int rism_autotune(rism_handle *h, int reps, int *geom_out) {
  if (!h || reps < 1) { set_error("rism_autotune: bad arguments"); return RISM_ERR_ARG; }
  if (!h->potentials) { set_error("rism_build_potentials must run before rism_autotune"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  const int rc = h->opts.precision == RISM_PREC_MIXED ? autotune_impl<float>(h, reps) : autotune_impl<double>(h, reps);
  if (rc) return rc;
  h->tuned = true;
  return geom_out ? rism_geometry(h, geom_out) : RISM_OK;
}

int rism_solve(rism_handle *h, const char *tuv_path, int *steps, double *rms) {
  if (!h) { set_error("null handle"); return RISM_ERR_ARG; }
  if (!h->potentials) { set_error("rism_build_potentials must run before rism_solve"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  if (!h->tuned && env_long("RISM_AUTOTUNE", 0) > 0)  // tune once per handle (cached per grid shape)
    if (int rc = rism_autotune(h, (int)env_long("RISM_AUTOTUNE", 0), nullptr)) return rc;
  return h->opts.precision == RISM_PREC_MIXED ? solve_impl<float>(h, tuv_path, steps, rms)
                                              : solve_impl<double>(h, tuv_path, steps, rms);
}

int rism_thermo(rism_handle *h, rism_xmu *out) {
  if (!h || !out) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->solved) { set_error("rism_solve must run before rism_thermo"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  const double t0 = now_s();
  if (!h->evaluated)
    if (int rc = evaluate(h)) return rc;
  int rc = thermo_impl<double>(h, out);
  if (rc == RISM_OK) rc = prof_collect(h);
  h->t_thermo = now_s() - t0;
  return rc;
}

int rism_grad(rism_handle *h, double *out) {
  if (!h || !out) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->solved) { set_error("rism_solve must run before rism_grad"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  if (!h->evaluated)
    if (int rc = evaluate(h)) return rc;
  const int nu = h->natu;
  if (nu == 0) return RISM_OK;
  const int nblk = (int)((h->N + (long)kThreads * 16 - 1) / ((long)kThreads * 16));
  const size_t bpart = align256((size_t)nblk * nu * 6 * sizeof(double));
  void *sp = nullptr;
  if (int rc = handle_scratch(h, bpart + (size_t)nu * 6 * sizeof(double), &sp)) return rc;
  double *part = (double *)sp, *dout = (double *)((char *)sp + bpart);
  hipError_t e = launch_grad(h->nx, h->ny, h->nz, h->dr, h->V, nu, h->atoms, h->lj, h->sigh, h->d_qv, h->d_rho,
                             h->gfield, 0, nu, part, nblk, dout, h->stream);
  std::vector<double> g((size_t)nu * 6);
  if (e == hipSuccess) e = hipMemcpyAsync(g.data(), dout, g.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_error(std::string("HIP error in rism_grad: ") + hipGetErrorString(e));
    return RISM_ERR_HIP;
  }
  // kernel layout per atom: LJ xyz then ES xyz, already x dv (output_grad.cu:15-24)
  std::memcpy(out, g.data(), g.size() * sizeof(double));
  return RISM_OK;
}

// rism_grad without any device allocation: the partial sums live in the
// library's static device array (RISM_NOALLOC_SCRATCH_BYTES per GPU, see
// grad_scratch_ptr below), the atoms taken in batches that fit it.  The
// kernels and their summation order are rism_grad's, so the results are
// identical.  Every partial and result slot is written before it is read,
// so nothing is cleared; calls that overlap on one GPU share the array and
// race.
int rism_grad_noalloc(rism_handle *h, double *out) {
  if (!h || !out) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->solved) { set_error("rism_solve must run before rism_grad_noalloc"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  const int nu = h->natu;
  if (nu == 0) return RISM_OK;
  const int nblk = (int)((h->N + (long)kThreads * 16 - 1) / ((long)kThreads * 16));
  const size_t bout = align256((size_t)nu * 6 * sizeof(double));
  const size_t per_atom = (size_t)nblk * 6 * sizeof(double);
  if (bout + per_atom > (size_t)RISM_NOALLOC_SCRATCH_BYTES) {
    set_error("rism_grad_noalloc: " + std::to_string(nu) + " atoms on " + std::to_string(h->N) +
              " points need " + std::to_string(bout + per_atom) + " bytes of scratch for one atom batch; the limit is " +
              std::to_string((long)RISM_NOALLOC_SCRATCH_BYTES));
    return RISM_ERR_ARG;
  }
  void *sp = nullptr;
  if (int rc = grad_scratch_ptr(h->device, &sp)) return rc;
  // evaluate() allocates nothing; per-kernel timing (which may create events) is off meanwhile
  const bool prof = h->prof;
  h->prof = false;
  int rc = h->evaluated ? RISM_OK : evaluate(h);
  h->prof = prof;
  if (rc) return rc;
  double *dout = (double *)sp, *part = (double *)((char *)sp + bout);
  const int nb = (int)std::min<size_t>((size_t)nu, ((size_t)RISM_NOALLOC_SCRATCH_BYTES - bout) / per_atom);
  hipError_t e = hipSuccess;
  for (int u0 = 0; u0 < nu && e == hipSuccess; u0 += nb)
    e = launch_grad(h->nx, h->ny, h->nz, h->dr, h->V, nu, h->atoms, h->lj, h->sigh, h->d_qv, h->d_rho, h->gfield,
                    u0, std::min(nb, nu - u0), part, nblk, dout, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)nu * 6 * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_error(std::string("HIP error in rism_grad_noalloc: ") + hipGetErrorString(e));
    return RISM_ERR_HIP;
  }
  return RISM_OK;
}

int rism_euv(rism_handle *h, double *out) {
  if (!h || !out) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->solved) { set_error("rism_solve must run before rism_euv"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  if (!h->evaluated)
    if (int rc = evaluate(h)) return rc;
  const int nu = h->natu, V = h->V;
  if (nu == 0) return RISM_OK;
  const int nblk = (int)((h->N + (long)kThreads * 16 - 1) / ((long)kThreads * 16));
  const size_t bpart = align256((size_t)nblk * nu * 2 * V * sizeof(double));
  void *sp = nullptr;
  if (int rc = handle_scratch(h, bpart + (size_t)nu * 2 * V * sizeof(double), &sp)) return rc;
  double *part = (double *)sp, *dout = (double *)((char *)sp + bpart);
  hipError_t e = launch_euv_atoms(h->nx, h->ny, h->nz, h->dr, V, nu, h->atoms, h->lj, h->sigh, h->d_qv, h->d_rho,
                                  h->gfield, part, nblk, dout, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, dout, (size_t)nu * 2 * V * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_error(std::string("HIP error in rism_euv: ") + hipGetErrorString(e));
    return RISM_ERR_HIP;
  }
  return RISM_OK;
}

int rism_qv(rism_handle *h, double *out) {
  if (!h || !out) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->solved) { set_error("rism_solve must run before rism_qv"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  if (!h->evaluated)
    if (int rc = evaluate(h)) return rc;
  void *sp = nullptr;
  if (int rc = handle_scratch(h, (size_t)h->N * sizeof(double), &sp)) return rc;
  double *d = (double *)sp;
  hipError_t e = launch_qv(h->N, h->V, h->d_rho, h->d_qv, h->gfield, d, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_error(std::string("HIP error in rism_qv: ") + hipGetErrorString(e));
    return RISM_ERR_HIP;
  }
  return RISM_OK;
}

// RMDFT (cal_rmdft.cu:5-114) on the FP64 evaluation of the final state:
// xmu1 = -sum_v rho_v sum h_v; w = IFFT(sum_v' c~_vv' rho_v' H_v');
// xmu2 = sum_v rho_v sum w_v, xmu3 = sum_v rho_v sum w_v h_v; the
// hard-sphere term eda from the weighted density; result (xmu1 + xmu2 +
// xmu3 / 2 - eda) dv RT.  Sums are per-block partials added in block order.
int rism_rmdft(rism_handle *h, const rism_rmdft_input *in, double *sfe) {
  if (!h || !in || !sfe || !in->cvk || !in->khs || !in->chs || !in->whs || !in->pfhs || !in->w0 || in->nhs < 4) {
    set_error("rism_rmdft: bad arguments");
    return RISM_ERR_ARG;
  }
  if (!h->solved) { set_error("rism_solve must run before rism_rmdft"); return RISM_ERR_STATE; }
  for (int i = 1; i < in->nhs; ++i)
    if (!(in->khs[i] > in->khs[i - 1])) { set_error("rism_rmdft: hard-sphere k table not ascending"); return RISM_ERR_ARG; }
  HIP_OK(hipSetDevice(h->device));
  if (!h->evaluated)
    if (int rc = evaluate(h)) return rc;
  const int V = h->V, nt = h->ntab, nh = in->nhs;
  const long N = h->N;
  // the reference stops when |k| runs beyond the tables (solvent_spline.cu:18-21, solvent_spline2.cu:16-19)
  const double kmax = std::sqrt(std::pow(M_PI * (h->nx - 1) / h->box[0], 2) + std::pow(M_PI * (h->ny - 1) / h->box[1], 2) +
                                std::pow(M_PI * (h->nz - 1) / h->box[2], 2));
  if (kmax > h->dk * nt || kmax > in->khs[nh - 1]) { set_error("insufficient maximal T tabulated"); return RISM_ERR_RANGE; }
  // spline coefficient tables (host, FP64)
  std::vector<double4> cvk((size_t)V * V * (nt - 1)), chs((size_t)V * (nh - 1)), whs((size_t)V * (nh - 1));
  {
    std::vector<double> kt(nt), y(std::max(nt, nh)), c0(std::max(nt, nh)), c1(std::max(nt, nh)), c2(std::max(nt, nh));
    for (int i = 0; i < nt; ++i) kt[i] = h->dk * (i + 1);
    for (int v = 0; v < V; ++v)
      for (int v1 = 0; v1 < V; ++v1) {
        for (int i = 0; i < nt; ++i) y[i] = in->cvk[((size_t)i * V + v) * V + v1];
        natural_spline(kt.data(), y.data(), nt, c0.data(), c1.data(), c2.data());
        for (int i = 0; i < nt - 1; ++i) cvk[(size_t)(v * V + v1) * (nt - 1) + i] = make_double4(y[i], c0[i], c1[i], c2[i]);
      }
    for (int v = 0; v < V; ++v)
      for (int which = 0; which < 2; ++which) {
        const double *src = (which ? in->whs : in->chs) + (size_t)v * nh;
        natural_spline(in->khs, src, nh, c0.data(), c1.data(), c2.data());
        std::vector<double4> &dst = which ? whs : chs;
        for (int i = 0; i < nh - 1; ++i) dst[(size_t)v * (nh - 1) + i] = make_double4(src[i], c0[i], c1[i], c2[i]);
      }
  }
  RmdftGrid g;
  g.nx = h->nx; g.ny = h->ny; g.nz = h->nz; g.N = N;
  for (int k = 0; k < 3; ++k) g.box[k] = h->box[k];
  RmdftTabs t;
  std::memset(&t, 0, sizeof t);
  t.V = V; t.ntab = nt; t.nhs = nh; t.dk = h->dk; t.rho0 = h->rho[0];
  for (int v = 0; v < V; ++v) t.rho[v] = h->rho[v];
  const int nblk = 1024;
  // device scratch: V transformed h, two work fields, tables, partials
  std::vector<void *> mem;
  auto grab = [&](size_t bytes) -> void * {
    void *p = nullptr;
    std::lock_guard<std::mutex> lk(g_device_mu);
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    mem.push_back(p);
    return p;
  };
  auto release = [&]() {
    std::lock_guard<std::mutex> lk(g_device_mu);
    for (void *p : mem) (void)hipFree(p);
    mem.clear();
  };
  std::vector<double2 *> H(V);
  bool ok = true;
  for (int v = 0; v < V && ok; ++v) ok = (H[v] = (double2 *)grab((size_t)N * sizeof(double2))) != nullptr;
  double2 *w1 = ok ? (double2 *)grab((size_t)N * sizeof(double2)) : nullptr;
  double2 *w2 = w1 ? (double2 *)grab((size_t)N * sizeof(double2)) : nullptr;
  double2 **dH = w2 ? (double2 **)grab(V * sizeof(double2 *)) : nullptr;
  double4 *dcvk = dH ? (double4 *)grab(cvk.size() * sizeof(double4)) : nullptr;
  double4 *dchs = dcvk ? (double4 *)grab(chs.size() * sizeof(double4)) : nullptr;
  double4 *dwhs = dchs ? (double4 *)grab(whs.size() * sizeof(double4)) : nullptr;
  double *dkhs = dwhs ? (double *)grab(nh * sizeof(double)) : nullptr;
  double *part = dkhs ? (double *)grab(2 * nblk * sizeof(double)) : nullptr;
  if (!part) {
    release();
    set_error("rism_rmdft: device allocation failed");
    return RISM_ERR_NOMEM;
  }
  t.cvk = dcvk; t.chs = dchs; t.whs = dwhs; t.khs = dkhs;
  const double2 *Z = (const double2 *)h->ps[0].Z;  // h of the FP64 evaluation, two sites per value
  std::vector<double> hp(2 * nblk);
  auto partial_sums = [&](double *s0, double *s1) -> hipError_t {
    hipError_t e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nblk; ++i) {
      a += hp[2 * i];
      b += hp[2 * i + 1];
    }
    *s0 = a;
    if (s1) *s1 = b;
    return e;
  };
  auto fft = [&](double2 *d, int dir) {
    return launch_fft3d<double>(d, h->ps[0].plan[0], h->ps[0].plan[1], h->ps[0].plan[2], h->ps[0].geom, h->nx, h->ny,
                                h->nz, h->dv, 1.0 / h->volume, dir, h->stream);
  };
  hipError_t e = hipMemcpyAsync(dH, H.data(), V * sizeof(double2 *), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dcvk, cvk.data(), cvk.size() * sizeof(double4), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dchs, chs.data(), chs.size() * sizeof(double4), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dwhs, whs.data(), whs.size() * sizeof(double4), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dkhs, in->khs, nh * sizeof(double), hipMemcpyHostToDevice, h->stream);
  double xmu1 = 0.0, xmu2 = 0.0, xmu3 = 0.0, eda = 0.0;
  for (int v = 0; v < V && e == hipSuccess; ++v) {
    e = launch_rm_unpack(Z, N, v, H[v], part, nblk, h->stream);  // one partial per block
    double s = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), part, nblk * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    for (int i = 0; i < nblk; ++i) s += hp[i];
    xmu1 -= s * h->rho[v];
    if (e == hipSuccess) e = fft(H[v], -1);
  }
  for (int v = 0; v < V && e == hipSuccess; ++v) {
    const double bex = (4.0 - 2.0 * in->pfhs[v]) / std::pow(1.0 - in->pfhs[v], 3) * (in->pfhs[v] / t.rho0);  // excp_fd.cc
    const double av = in->pfhs[v] / t.rho0;
    const double e0 = t.rho0 * bex * in->w0[v];
    double s0 = 0.0, s1 = 0.0;
    e = launch_rm_kmix(g, t, v, dH, w1, w2, h->stream);
    if (e == hipSuccess) e = fft(w1, 1);
    if (e == hipSuccess) e = launch_rm_reduce(Z, N, v, w1, 0, 0.0, part, nblk, h->stream);
    if (e == hipSuccess) e = partial_sums(&s0, &s1);
    xmu2 += s0 * h->rho[v];
    xmu3 += s1 * h->rho[v];
    if (e == hipSuccess) e = fft(w2, 1);
    if (e == hipSuccess) e = launch_rm_fd(Z, N, v, av, bex, t.rho0, w2, h->stream);
    if (e == hipSuccess) e = fft(w2, -1);
    if (e == hipSuccess) e = launch_rm_wmul(g, t, v, w2, h->stream);
    if (e == hipSuccess) e = fft(w2, 1);
    if (e == hipSuccess) e = launch_rm_reduce(Z, N, v, w2, 1, e0, part, nblk, h->stream);
    if (e == hipSuccess) e = partial_sums(&s0, nullptr);
    eda += s0 * h->rho[v];
  }
  release();
  if (e != hipSuccess) {
    set_error(std::string("HIP error in rism_rmdft: ") + hipGetErrorString(e));
    return RISM_ERR_HIP;
  }
  *sfe = kAvogadro * kBoltzmann * h->temper * (xmu1 + xmu2 + 0.5 * xmu3 - eda) * h->dv;  // output_xmu.cu:25-28
  return RISM_OK;
}

// Adjoint of the converged fixed point (FP64 in both precisions); see
// k_adjoint.hip for the equations.  The adjoint system
//   lambda = (g' - 1)(K^T lambda - lambda) + b
// is the transpose of the forward map's linearisation and is solved with
// the same modified Anderson scheme, on its own control block, to
// rms(residual) <= rtol * rms(b).
int rism_adjoint_grad(rism_handle *h, double rtol, double *dxyz, double *dq, double *dsig, double *deps,
                      int *steps) {
  if (!h) { set_error("null handle"); return RISM_ERR_ARG; }
  if (!h->solved) { set_error("rism_solve must run before rism_adjoint_grad"); return RISM_ERR_STATE; }
  if (!(rtol > 0.0)) { set_error("rtol must be positive"); return RISM_ERR_ARG; }
  HIP_OK(hipSetDevice(h->device));
  if (!h->evaluated)
    if (int rc = evaluate(h)) return rc;
  const long N = h->N;
  const int V = h->V, P = h->P, nu = h->natu, nh = h->opts.nhist;
  const size_t VN = (size_t)V * N;
  PrecSet &ps = h->ps[0];
  int rc = 0;
  if (!h->adj_f) {  // grid-sized: kept for the life of the handle
    if ((rc = dalloc_t(h, &h->adj_f, 7 * VN)) || (rc = dalloc_t(h, &h->adj_hist, 2 * (size_t)nh * VN)) ||
        (rc = dalloc_t(h, &h->adj_n, 2 * (size_t)N)) || (rc = dalloc_t(h, &h->adj_Z, (size_t)P * N)) ||
        (rc = dalloc_t(h, &h->adj_Y, (size_t)N)) || (rc = dalloc_t(h, &h->adj_rms, (size_t)h->opts.maxstep + 4)) ||
        (rc = dalloc_t(h, &h->adj_ctl, 1)))
      return rc;
  }
  if (h->adj_ljd_cap < std::max(1, nu)) {  // sized by the solute
    dfree(h, h->adj_ljd);
    h->adj_ljd = nullptr;
    h->adj_ljd_cap = 0;
    if ((rc = dalloc_t(h, &h->adj_ljd, (size_t)V * std::max(1, nu)))) return rc;
    h->adj_ljd_cap = std::max(1, nu);
  }
  // (1 / (2 sigma_uv), 4 eps_uv / (2 eps_u)): Lorentz-Berthelot derivatives
  std::vector<double2> ljd((size_t)V * std::max(1, nu));
  for (int v = 0; v < V; ++v)
    for (int u = 0; u < nu; ++u) {
      const double s = (h->usig[u] + h->sigv[v]) * 0.5;
      const double e = std::sqrt(h->ueps[u] * h->epsv[v]);
      ljd[(size_t)v * nu + u] = make_double2(s > 0.0 ? 0.5 / s : 0.0, h->ueps[u] > 0.0 ? 2.0 * e / h->ueps[u] : 0.0);
    }
  HIP_OK(hipMemcpyAsync(h->adj_ljd, ljd.data(), ljd.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  AdjArgs a;
  std::memset(&a, 0, sizeof a);
  a.N = N; a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.V = V; a.closure = h->opts.closure; a.natu = nu;
  a.ntab = h->ntab;
  for (int k = 0; k < 3; ++k) { a.box[k] = h->box[k]; a.dr[k] = h->dr[k]; }
  a.dv = h->dv; a.dk = h->dk;
  a.kt = kAvogadro * kBoltzmann * h->temper;
  a.ikbt = 1.0 / a.kt;
  a.ubeta = kHartree * kBohr / (kBoltzmann * h->temper);
  for (int v = 0; v < V; ++v) {
    a.qcf[v] = h->qv[v] * h->cf;
    a.qv[v] = h->qv[v];
    a.rho[v] = h->rho[v];
    a.rho_inv[v] = 1.0 / h->rho[v];
  }
  a.ts = h->ts; a.u = (const double *)ps.u; a.fr = h->fr;
  a.hz = (const double2 *)ps.Z; a.chi = (const double4 *)ps.chi;
  a.gp = h->adj_f; a.alpha = a.gp + VN; a.beta = a.alpha + VN; a.b = a.beta + VN;
  a.lam = a.b + VN; a.res = a.lam + VN; a.ubar = a.res + VN;
  a.hist_l = h->adj_hist; a.hist_r = h->adj_hist + (size_t)nh * VN;
  a.debar = h->adj_n; a.frbar = h->adj_n + N;
  a.ctl = h->adj_ctl; a.partials = h->partials;
  a.atoms = h->atoms; a.lj = h->lj; a.ljd = h->adj_ljd; a.phase = h->phase;
  hipStream_t st = h->stream;
  const double ivol = 1.0 / h->volume;
  auto fft = [&](int dir) -> hipError_t {
    for (int p = 0; p < P; ++p) {
      hipError_t e = launch_fft3d<double>(h->adj_Z + (size_t)p * N, ps.plan[0], ps.plan[1], ps.plan[2], ps.geom,
                                          h->nx, h->ny, h->nz, h->dv, ivol, dir, st);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  // apply K^T to x (+ y) into adj_Z (real space, packed); Y receives the k-space weight field
  auto apply_kt = [&](const double *x, const double *y, double2 *Y) -> hipError_t {
    hipError_t e = launch_adj_pack(a, x, y, h->adj_Z, st);
    if (e == hipSuccess) e = fft(-1);
    if (e == hipSuccess) e = launch_adj_kspace(a, h->adj_Z, Y, st);
    if (e == hipSuccess) e = fft(+1);
    return e;
  };
  // The mixed path runs the iterations' K^T applications with FP32
  // transforms on its FP32 spectrum buffer (idle after the solve); the right
  // side b and the final application behind the field adjoints stay FP64.
  const bool f32 = h->opts.precision == RISM_PREC_MIXED;
  float2 *Z32 = f32 ? (float2 *)h->ps[1].Z : nullptr;
  auto apply_kt32 = [&](const double *x) -> hipError_t {
    const PrecSet &p1 = h->ps[1];
    hipError_t e = launch_adj_pack(a, x, (const double *)nullptr, Z32, st);
    for (int p = 0; p < P && e == hipSuccess; ++p)
      e = launch_fft3d<float>(Z32 + (size_t)p * N, p1.plan[0], p1.plan[1], p1.plan[2], p1.geom, h->nx, h->ny, h->nz,
                              h->dv, ivol, -1, st);
    if (e == hipSuccess) e = launch_adj_kspace(a, Z32, (double2 *)nullptr, st);
    for (int p = 0; p < P && e == hipSuccess; ++p)
      e = launch_fft3d<float>(Z32 + (size_t)p * N, p1.plan[0], p1.plan[1], p1.plan[2], p1.geom, h->nx, h->ny, h->nz,
                              h->dv, ivol, +1, st);
    return e;
  };
  HIP_OK(launch_adj_fields(a, st));
  HIP_OK(apply_kt(a.alpha, nullptr, nullptr));
  HIP_OK(launch_adj_rhs(a, h->adj_Z, st));
  HIP_OK(hipMemsetAsync(a.lam, 0, VN * sizeof(double), st));
  HIP_OK(hipMemsetAsync(h->adj_hist, 0, 2 * (size_t)nh * VN * sizeof(double), st));
  Ctl c;
  std::memset(&c, 0, sizeof c);
  c.run_iter = 0; c.upd_iter = -1; c.status = -1;
  c.maxstep = h->opts.maxstep; c.count = h->opts.count; c.nhist = nh;
  c.tol = -1.0;  // set after the first step from rms(b)
  c.m = h->opts.m; c.mp = h->opts.mp; c.npts = (double)VN;
  HIP_OK(hipMemcpyAsync(h->adj_ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
  // FP32 transforms carry about 1e-7 of relative rounding: the mixed path's
  // adjoint solve stops at rms <= max(rtol, 1e-6) rms(b)
  if (f32) rtol = std::max(rtol, 1e-6);
  const int nblk = ps.geom.xblocks;
  const int nv = 1 + h->hm + h->hm * (h->hm + 1) / 2;
  int it = 0;
  for (;;) {
    HIP_OK(launch_adj_update(a, h->hm, st));
    if (f32) {
      HIP_OK(apply_kt32(a.lam));
      HIP_OK(launch_adj_resid(a, (const float2 *)Z32, h->hm, nblk, st));
    } else {
      HIP_OK(apply_kt(a.lam, nullptr, nullptr));
      HIP_OK(launch_adj_resid(a, (const double2 *)h->adj_Z, h->hm, nblk, st));
    }
    HIP_OK(launch_anderson_solve(h->adj_ctl, h->partials, nblk, nv, h->hm, h->adj_rms, st));
    ++it;
    if (it == 1 || it % 8 == 0 || it >= h->opts.maxstep + 1) {
      HIP_OK(hipMemcpyAsync(&c, h->adj_ctl, sizeof c, hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      if (it == 1 && c.status == -1) {  // lambda = 0: the first residual is b
        c.tol = rtol * c.rms;
        HIP_OK(hipMemcpyAsync(h->adj_ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
      }
      if (c.status != -1 || it >= h->opts.maxstep + 1) break;
    }
  }
  if (c.status == RISM_SOLVE_MAXSTEP && c.upd_hold == c.cur) {
    // a step-limit stop: its pending update, held by the solve kernel, once
    c.upd_iter = c.cur;
    c.run_iter = -1;
    HIP_OK(hipMemcpyAsync(h->adj_ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
    HIP_OK(launch_adj_update(a, h->hm, st));
    HIP_OK(hipStreamSynchronize(st));
  }
  if (steps) *steps = c.total;
  if (c.status == RISM_SOLVE_DIVERGED || c.status == -1) {
    set_error("adjoint solve did not converge");
    return RISM_SOLVE_DIVERGED;
  }
  // field adjoints from mu = lambda + alpha, and the k-space weight Y
  HIP_OK(apply_kt(a.lam, a.alpha, h->adj_Y));
  HIP_OK(launch_adj_bar(a, h->adj_Z, st));
  if (nu > 0) {
    const int gblk = adj_grad_blocks(N);
    const size_t b6 = align256((size_t)gblk * nu * 6 * sizeof(double)),
                 b4 = align256((size_t)gblk * nu * 4 * sizeof(double));
    void *sp = nullptr;
    if (int rc2 = handle_scratch(h, b6 + b4 + (size_t)nu * 10 * sizeof(double), &sp)) return rc2;
    double *p6 = (double *)sp, *p4 = (double *)((char *)sp + b6), *o = (double *)((char *)sp + b6 + b4);
    hipError_t e = launch_adj_grad(a, h->adj_Y, p6, p4, o, o + (size_t)nu * 6, st);
    std::vector<double> r((size_t)nu * 10);
    if (e == hipSuccess) e = hipMemcpyAsync(r.data(), o, r.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error(std::string("HIP error in rism_adjoint_grad: ") + hipGetErrorString(e));
      return RISM_ERR_HIP;
    }
    const double *r6 = r.data(), *r4 = r.data() + (size_t)nu * 6;
    const double ks = -h->cf * a.ubeta / (h->volume * h->dv);  // fk term: -(cf / (vol dv)) sum_k Y dfk
    for (int u = 0; u < nu; ++u) {
      if (dxyz)
        for (int k = 0; k < 3; ++k) dxyz[3 * u + k] = r6[6 * u + k] + ks * h->uq[u] * r4[4 * u + 1 + k];
      if (dq) dq[u] = r6[6 * u + 3] + ks * r4[4 * u];
      if (dsig) dsig[u] = r6[6 * u + 4];
      if (deps) deps[u] = r6[6 * u + 5];
    }
  }
  HIP_OK(hipStreamSynchronize(st));
  return c.status == 0 ? RISM_OK : c.status;
}

int rism_get_field(rism_handle *h, int which, double *out) {
  if (!h || !out) { set_error("null argument"); return RISM_ERR_ARG; }
  HIP_OK(hipSetDevice(h->device));
  HIP_OK(hipStreamSynchronize(h->stream));
  const long N = h->N;
  const int V = h->V;
  const size_t VN = (size_t)V * N;
  switch (which) {
    case RISM_FIELD_T: return fetch_t(h, out);
    case RISM_FIELD_RES:
      if (h->opts.precision != RISM_PREC_MIXED) {
        HIP_OK(hcopy(h, out, h->res, VN * sizeof(double), hipMemcpyDeviceToHost));
      } else {  // the last residual pass wrote ring slot rlast
        if (int rc = get_ctl(h)) return rc;
        std::vector<float> r(VN);
        HIP_OK(hcopy(h, r.data(), h->mring + (size_t)h->ctl_host->rlast * VN, VN * sizeof(float),
                     hipMemcpyDeviceToHost));
        for (size_t i = 0; i < VN; ++i) out[i] = r[i];
      }
      return RISM_OK;
    case RISM_FIELD_ULJ: HIP_OK(hcopy(h, out, h->ulj, VN * sizeof(double), hipMemcpyDeviceToHost)); return RISM_OK;
    case RISM_FIELD_DE: HIP_OK(hcopy(h, out, h->de, N * sizeof(double), hipMemcpyDeviceToHost)); return RISM_OK;
    case RISM_FIELD_FR: HIP_OK(hcopy(h, out, h->fr, N * sizeof(double), hipMemcpyDeviceToHost)); return RISM_OK;
    case RISM_FIELD_FK: {
      const int nx = h->nx, ny = h->ny, nz = h->nz, nyh = ny / 2;
      std::vector<double2> half((size_t)nx * nyh * nz);
      HIP_OK(hcopy(h, half.data(), h->fkh, half.size() * sizeof(double2), hipMemcpyDeviceToHost));
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            const size_t ip = ((size_t)z * ny + y) * nx + x;
            double2 v;
            if (y < nyh) {
              v = half[((size_t)z * nyh + y) * nx + x];
            } else {  // fk(-k) = conj(fk(k))
              const double2 m = half[((size_t)(nz - 1 - z) * nyh + (ny - 1 - y)) * nx + (nx - 1 - x)];
              v = make_double2(m.x, -m.y);
            }
            out[2 * ip] = v.x;
            out[2 * ip + 1] = v.y;
          }
      return RISM_OK;
    }
    case RISM_FIELD_CHI: {
      const int nt = h->ntab;
      std::vector<double> kt(nt), y(nt), c0(nt), c1(nt), c2(nt);
      for (int i = 0; i < nt; ++i) kt[i] = h->dk * (i + 1);
      for (int v = 0; v < V; ++v)
        for (int v1 = 0; v1 < V; ++v1) {
          for (int i = 0; i < nt; ++i) y[i] = h->chi[((size_t)i * V + v) * V + v1];
          natural_spline(kt.data(), y.data(), nt, c0.data(), c1.data(), c2.data());
          double *o = out + (size_t)(v * V + v1) * N;
          for (long ip = 0; ip < N; ++ip) {
            const int x = (int)(ip % h->nx), yy = (int)((ip / h->nx) % h->ny), z = (int)(ip / ((long)h->nx * h->ny));
            const double kx = 2.0 * M_PI * ((double)x - h->nx / 2.0 + 0.5) / h->box[0];
            const double ky = 2.0 * M_PI * ((double)yy - h->ny / 2.0 + 0.5) / h->box[1];
            const double kz = 2.0 * M_PI * ((double)z - h->nz / 2.0 + 0.5) / h->box[2];
            const double kk = std::sqrt(kx * kx + ky * ky + kz * kz);
            const int lo = table_interval(kk, h->dk, nt);
            const double t = kk - h->dk * (lo + 1);
            o[ip] = ((c2[lo] * t + c1[lo]) * t + c0[lo]) * t + y[lo];
          }
        }
      return RISM_OK;
    }
    case RISM_FIELD_G:
    case RISM_FIELD_H:
    case RISM_FIELD_C: {
      if (!h->solved) { set_error("fields g/h/c need a solve"); return RISM_ERR_STATE; }
      if (!h->evaluated)
        if (int rc = evaluate(h)) return rc;
      if (which == RISM_FIELD_G) {
        HIP_OK(hcopy(h, out, h->gfield, VN * sizeof(double), hipMemcpyDeviceToHost));
        return RISM_OK;
      }
      std::vector<double> hh(VN);
      {
        std::vector<double2> z((size_t)h->P * N);  // FP64 evaluation result
        HIP_OK(hcopy(h, z.data(), h->ps[0].Z, z.size() * sizeof(double2), hipMemcpyDeviceToHost));
        for (int v = 0; v < V; ++v)
          for (long i = 0; i < N; ++i) hh[(size_t)v * N + i] = (v & 1) ? z[(size_t)(v >> 1) * N + i].y : z[(size_t)(v >> 1) * N + i].x;
      }
      if (which == RISM_FIELD_H) {
        std::memcpy(out, hh.data(), VN * sizeof(double));
        return RISM_OK;
      }
      if (int rc = fetch_t(h, out)) return rc;
      for (size_t i = 0; i < VN; ++i) out[i] = hh[i] - out[i];
      return RISM_OK;
    }
    default:
      set_error("unknown field");
      return RISM_ERR_ARG;
  }
}

int rism_set_t(rism_handle *h, const double *t) {
  if (!h || !t) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->potentials) { set_error("potentials needed before rism_set_t"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  h->evaluated = false;
  return store_t(h, t);
}

int rism_save_tuv(rism_handle *h, const char *path) {
  if (!h || !path) { set_error("null argument"); return RISM_ERR_ARG; }
  HIP_OK(hipSetDevice(h->device));
  return write_tuv_file(h, path);
}

int rism_load_tuv(rism_handle *h, const char *path) {
  if (!h || !path) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->potentials) { set_error("potentials needed before rism_load_tuv"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  h->cf = 1.0;
  h->evaluated = false;
  return read_tuv_file(h, path);
}

int rism_stats(rism_handle *h, int *steps, double *rms, double *t_setup, double *t_solve, double *t_thermo) {
  if (!h) { set_error("null handle"); return RISM_ERR_ARG; }
  if (steps) *steps = h->steps;
  if (rms) *rms = h->rms;
  if (t_setup) *t_setup = h->t_setup;
  if (t_solve) *t_solve = h->t_solve;
  if (t_thermo) *t_thermo = h->t_thermo;
  return RISM_OK;
}

int rism_rms_history(rism_handle *h, double *out, int n) {
  if (!h) { set_error("null handle"); return RISM_ERR_ARG; }
  const int k = std::min(n, (int)h->rms_log.size());
  for (int i = 0; i < k; ++i) out[i] = h->rms_log[i];
  return (int)h->rms_log.size();
}

void rism_destroy(rism_handle *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  std::lock_guard<std::mutex> lk(g_device_mu);
  drop_graphs(h);
  for (auto &b : h->bufs) release_locked(h->device, b);
  for (auto &e : h->ev_pool) (void)hipEventDestroy(e);
  if (h->borrowed_ctx) {
    // the batch's stream and pinned block: its owner releases them
  } else if (h->stream && h->ctl_host && g_spare.size() < kMaxSpare) {
    g_spare.push_back({h->device, h->stream, h->ctl_host});
  } else {
    if (h->ctl_host) (void)hipHostFree(h->ctl_host);
    if (h->stream) (void)hipStreamDestroy(h->stream);
  }
  delete h;
}

long rism_release_cache(void) {
  std::lock_guard<std::mutex> lk(g_device_mu);
  const long freed = (long)g_cache_bytes;
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &c : g_cache) {
    (void)hipSetDevice(c.device);
    (void)hipFree(c.p);
  }
  g_cache.clear();
  g_cache_bytes = 0;
  for (auto &sp : g_spare) {
    (void)hipSetDevice(sp.device);
    (void)hipHostFree(sp.ctl_host);
    (void)hipStreamDestroy(sp.stream);
  }
  g_spare.clear();
  (void)hipSetDevice(cur);
  return freed;
}

int rism_profile(rism_handle *h, int enable) {
  if (!h) { set_error("null handle"); return RISM_ERR_ARG; }
  HIP_OK(hipSetDevice(h->device));
  if (int rc = prof_collect(h)) return rc;
  h->prof = enable != 0;
  for (int k = 0; k < RISM_NKERNEL; ++k) { h->kt_ms[k] = 0.0; h->kt_n[k] = 0; }
  return RISM_OK;
}

int rism_kernel_times(rism_handle *h, double *ms, long *count) {
  if (!h) { set_error("null handle"); return RISM_ERR_ARG; }
  if (int rc = prof_collect(h)) return rc;
  for (int k = 0; k < RISM_NKERNEL; ++k) {
    if (ms) ms[k] = h->kt_ms[k];
    if (count) count[k] = h->kt_n[k];
  }
  return RISM_OK;
}

int rism_geometry(rism_handle *h, int *out) {
  if (!h || !out) { set_error("null argument"); return RISM_ERR_ARG; }
  const LaunchGeom &g = h->ps[h->opts.precision == RISM_PREC_MIXED ? 1 : 0].geom;
  const int v[10] = {g.Lx, g.Ty, g.Tz, g.ept_x, g.ept_y, g.ept_z, g.xgroups, g.xblocks, h->P, h->hm};
  for (int k = 0; k < 10; ++k) out[k] = v[k];
  return RISM_OK;
}

int rism_fft3d(rism_handle *h, void *data, int dir) {
  if (!h || !data || (dir != -1 && dir != 1)) { set_error("bad argument"); return RISM_ERR_ARG; }
  HIP_OK(hipSetDevice(h->device));
  hipError_t e;
  if (h->opts.precision == RISM_PREC_MIXED)
    e = launch_fft3d<float>((float2 *)data, h->ps[1].plan[0], h->ps[1].plan[1], h->ps[1].plan[2], h->ps[1].geom, h->nx, h->ny, h->nz,
                            h->dv, 1.0 / h->volume, dir, h->stream);
  else
    e = launch_fft3d<double>((double2 *)data, h->ps[0].plan[0], h->ps[0].plan[1], h->ps[0].plan[2], h->ps[0].geom, h->nx, h->ny, h->nz,
                             h->dv, 1.0 / h->volume, dir, h->stream);
  HIP_OK(e);
  HIP_OK(hipStreamSynchronize(h->stream));
  return RISM_OK;
}

int rism_calculate(rism_handle *h, const double *t_dev, double *res_dev) {
  if (!h || !t_dev || !res_dev) { set_error("null argument"); return RISM_ERR_ARG; }
  if (!h->potentials) { set_error("potentials needed before rism_calculate"); return RISM_ERR_STATE; }
  HIP_OK(hipSetDevice(h->device));
  const size_t VN = (size_t)h->V * h->N;
  h->cf = 1.0;
  // ts = t - q_v fr
  HIP_OK(hipMemcpyAsync(h->ts, t_dev, VN * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  std::vector<double> coef(h->V);
  for (int v = 0; v < h->V; ++v) coef[v] = -h->qv[v];
  HIP_OK(launch_axpy_fr<double>(h->ts, h->fr, coef.data(), h->V, h->N, 1, h->stream));
  int rc = stage_potentials(h);
  if (rc) return rc;
  Ctl &c = *h->ctl_host;
  std::memset(&c, 0, sizeof c);
  c.nhist = h->opts.nhist;
  c.maxstep = 1;
  c.npts = (double)VN;
  c.status = -1;
  c.run_iter = c.cur;
  c.upd_iter = -1;
  c.refresh_iter = -1;
  const bool mixed = h->opts.precision == RISM_PREC_MIXED;
  if (mixed) {  // the FP64 refresh chain is the map itself (R rounded to FP32)
    c.ring = 1;
    c.rcur = h->opts.nhist;
    for (int k = 0; k < h->opts.nhist; ++k) c.rmap[k] = k;
    c.refresh_iter = c.cur;
    if ((rc = mx_convert(h, 0))) return rc;
  }
  if ((rc = put_ctl(h))) return rc;
  rc = mixed ? launch_mx_iteration(h, 2, false) : launch_iteration<double>(h, true, false);
  if (rc) return rc;
  if (!mixed)
    HIP_OK(hipMemcpyAsync(res_dev, h->res, VN * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  else
    HIP_OK(launch_widen(h->mring + (size_t)c.rcur * VN, res_dev, (long)VN, h->stream));
  h->mx_live = mixed;
  HIP_OK(hipStreamSynchronize(h->stream));
  h->solved = true;
  h->evaluated = false;
  return RISM_OK;
}

}  // extern "C"
