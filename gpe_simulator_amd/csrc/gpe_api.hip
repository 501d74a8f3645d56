// C ABI of libgpe_hip.so: contexts, solves, step orchestration and
// diagnostics.  Declarations and reference citations: include/gpe_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "gpe_hip.h"
#include "gpe_jit.h"
#include "gpe_misc.h"

using namespace gpe;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(GPE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(expr)                  \
  do {                             \
    int r_ = (expr);               \
    if (r_ != GPE_OK) return r_;   \
  } while (0)

constexpr int kPassKinds = 21;   // GPE_PASS_KINDS: the kinds a step uses (timed and stamped)
// The plane plan of a solve with the momentum cutoff (internal step plan: the
// API calls it GPE_PLAN_PLANE).
constexpr int kPlanPlaneCut = 3;
// Bytes of time-dependent slot tables per program: a 256^3 complex64 volume
// program (64 MiB a slot) evaluates 16 slots per 15 steps within it; with 64
// MiB it evaluated 2 slots every step, half of them again the next step.
constexpr long long kDynBudget = 1LL << 30;
constexpr long long kStageElems = 1LL << 20;   // complex128 elements per staging copy
const double kUnitScale = 1.0;                  // host source of unit pending scales

struct DevProgram {
  int* code = nullptr;
  double* consts = nullptr;
  int ncode = 0;
  bool active = false;
  JitProgram* jit = nullptr;     // run-time compiled tables (time-dependent programs); null: the interpreter
  double* hv = nullptr;          // the compiled program's hoisted values, kMaxSlots rows
};
constexpr int kMaxSlots = 258;   // time slots of one table evaluation: a chunk of at most 256 steps, plus one

// Time-dependent programs registered from now on run compiled (1, the
// default) or through the interpreter (0).
std::atomic<int> g_jit_enabled{1};

struct PassArgsRaw {
  const void* s;
  int slot_a, slot_b;
  int flags;
  int hist_index;
  int stamp_index;
  int norm_parts;
  double* red;
};
constexpr long long kMaxStamps = 1LL << 22;  // wall-clock slots of one stamped call (32 MiB)
static_assert(sizeof(PassArgsRaw) == sizeof(PassArgs<float>), "pass argument layout");
static_assert(sizeof(PassArgsRaw) == sizeof(PassArgs<double>), "pass argument layout");

int nbins_for(int n) { return int(std::floor(std::sqrt(3.0) * (n / 2) + 0.5)) + 1; }
// Bins of the equal-width view kept after the reference's shells.
constexpr int kViewBins = GPE_VIEW_BINS;

// Compiled pass variants.  Entries of one (precision, side) are contiguous and
// ordered by variant index; sides without a fast transform use the generic
// entry of their precision.
struct PassVariant {
  int prec, n, v, tp, b, waves;
  PassLauncher launch;
  PassBlocks blocks;
};

const PassVariant kVariantTable[] = {
    {0, 0, 0, 64, 1, 1, launch_pass_p0_n0_v0, pass_blocks_p0_n0_v0},
    {1, 0, 0, 64, 1, 1, launch_pass_p1_n0_v0, pass_blocks_p1_n0_v0},
#define GPE_VARIANT_ENTRY(P, N, V, TP, B, W, RD) \
  {P, N, V, TP, B, W, launch_pass_p##P##_n##N##_v##V, pass_blocks_p##P##_n##N##_v##V},
    GPE_VARIANTS(GPE_VARIANT_ENTRY)
#undef GPE_VARIANT_ENTRY
};

// First variant and count for (precision, side).
int variants_for(int prec, int n, const PassVariant** first) {
  const int total = int(sizeof(kVariantTable) / sizeof(kVariantTable[0]));
  int count = 0;
  for (int i = 0; i < total; ++i) {
    if (kVariantTable[i].prec == prec && kVariantTable[i].n == n) {
      if (count == 0) *first = &kVariantTable[i];
      ++count;
    }
  }
  if (count == 0) {
    *first = &kVariantTable[prec == 0 ? 0 : 1];
    count = 1;
  }
  return count;
}

// Per-pass timing with the semantics of a kernel trace: each timed launch
// hands its dispatch a stop event (hipExtLaunchKernel timestamps it at the
// dispatch's end), and a pass lasts from the previous timed pass's stop to
// its own, so back-to-back passes are timed as they run with nothing added
// between them.  When other work was enqueued since the last timed pass (the
// first pass of a call, a time-dependent table evaluation), a start marker is
// recorded before the launch instead.  Events come from a pool that grows on
// use and are read after the last launch; nothing synchronises in between.
struct PassTimer {
  struct Span {
    int kind, start, stop;     // pool indices
  };
  std::vector<hipEvent_t> pool;
  std::vector<Span> spans;
  size_t used = 0;
  bool chained = false;        // the stream's last command was a timed pass
  // own = true: each launch carries its own start and stop event, both
  // timestamped by the dispatch itself, so a span is the kernel's execution
  // alone (the figure a kernel trace reports), with nothing recorded between
  // launches
  bool own = false;
  double ms[kPassKinds] = {0};
  long long cnt[kPassKinds] = {0};

  void reset() {
    spans.clear();
    used = 0;
    chained = false;
    own = false;
    for (int k = 0; k < kPassKinds; ++k) {
      ms[k] = 0;
      cnt[k] = 0;
    }
  }
  hipError_t take(int* idx) {
    if (used == pool.size()) {
      hipEvent_t e = nullptr;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      pool.push_back(e);
    }
    *idx = (int)used++;
    return hipSuccess;
  }
  // Before a timed launch of `kind`: the events to hand to the launch (the
  // start event only in `own` mode; otherwise the span starts at the previous
  // timed launch's stop, or at a marker recorded here).
  hipError_t begin(int kind, hipStream_t st, hipEvent_t* start, hipEvent_t* stop) {
    int a = !own && chained ? spans.back().stop : -1, b = -1;
    hipError_t r = hipSuccess;
    if (a < 0 && (r = take(&a)) == hipSuccess && !own) r = hipEventRecord(pool[a], st);
    if (r == hipSuccess) r = take(&b);
    if (r != hipSuccess) return r;
    spans.push_back({kind, a, b});
    chained = true;
    *start = own ? pool[a] : nullptr;
    *stop = pool[b];
    return hipSuccess;
  }
  // Waits for the stream and adds every span to ms/cnt.
  hipError_t collect(hipStream_t st) {
    hipError_t r = hipStreamSynchronize(st);
    for (size_t i = 0; r == hipSuccess && i < spans.size(); ++i) {
      float t = 0.f;
      r = hipEventElapsedTime(&t, pool[spans[i].start], pool[spans[i].stop]);
      ms[spans[i].kind] += t;
      cnt[spans[i].kind] += 1;
    }
    spans.clear();
    used = 0;
    chained = false;
    return r;
  }
  void destroy() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

}  // namespace

// ------------------------------------------------- allocation-free scratch --
// Device memory of gpe_step_noalloc: the FP64 norm partials of an
// imaginary-time step (one per Z_LAST block: n^2 / B for a fast side, n^2 for
// the generic transform), so that the call allocates nothing.  Each GPU the
// library runs on holds one copy, at its own address, until the process ends.
// In host code `scratch` names a host-side stand-in, never the device memory:
// its per-device address comes from hipGetSymbolAddress once per device (at
// gpe_ctx_create, with the device current) and is the only handle used.
constexpr long long kScratchDoubles = 1LL << 16;   // 512 KiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
double* g_scratch_ptr[kMaxDevices] = {nullptr};   // per-device address of `scratch`
}  // namespace

struct gpe_ctx {
  int device = 0;
  double* scratch = nullptr;              // this device's `scratch` (device address)
  // per-pass-kind variant defaults for new solves, keyed by (side, precision, kind)
  std::map<std::tuple<int, int, int>, std::array<int, kPassKinds>> variants;
  // streams and pinned time-list buffers of destroyed solves and batches,
  // handed to the next ones (creating and destroying them costs about as
  // much as the rest of a solve's set-up)
  std::mutex spare_lock;
  std::vector<hipStream_t> spare_streams;
  std::multimap<size_t, void*> spare_pinned;
};

namespace {
constexpr size_t kSpareStreams = 256, kSparePinned = 512;

hipError_t take_stream(gpe_ctx* ctx, hipStream_t* st) {
  {
    std::lock_guard<std::mutex> guard(ctx->spare_lock);
    if (!ctx->spare_streams.empty()) {
      *st = ctx->spare_streams.back();
      ctx->spare_streams.pop_back();
      return hipSuccess;
    }
  }
  return hipStreamCreateWithFlags(st, hipStreamNonBlocking);
}

// (the stream has finished its work)
void give_stream(gpe_ctx* ctx, hipStream_t st) {
  if (!st) return;
  {
    std::lock_guard<std::mutex> guard(ctx->spare_lock);
    if (ctx->spare_streams.size() < kSpareStreams) {
      ctx->spare_streams.push_back(st);
      return;
    }
  }
  (void)hipStreamDestroy(st);
}

hipError_t take_pinned(gpe_ctx* ctx, size_t bytes, void** p) {
  {
    std::lock_guard<std::mutex> guard(ctx->spare_lock);
    auto it = ctx->spare_pinned.find(bytes);
    if (it != ctx->spare_pinned.end()) {
      *p = it->second;
      ctx->spare_pinned.erase(it);
      return hipSuccess;
    }
  }
  return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

// (no copy from the buffer is in flight)
void give_pinned(gpe_ctx* ctx, size_t bytes, void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> guard(ctx->spare_lock);
    if (ctx->spare_pinned.size() < kSparePinned) {
      ctx->spare_pinned.emplace(bytes, p);
      return;
    }
  }
  (void)hipHostFree(p);
}
}  // namespace

// Pass kinds outside the tuned set (the reference-structured step) run the
// variant chosen for the tuned kind along the same axis.
int tuned_kind(int kind) {
  if (kind < kPassKinds) return kind;
  return kind == K_Z_HALF || kind == K_Z_INV ? K_Z_FIRST : K_X_KIN;
}

struct gpe_solve {
  gpe_ctx* ctx = nullptr;
  int n = 0, prec = 0, kind = 0, masked = 0, nbins = 0;
  bool view_ready = false;                // a gpe_diagnostics call filled the view
  double dt = 0, beta = 0, kcut = 0, coef = 1, dx3 = 0;
  long long n3 = 0;
  size_t esz = 0, rsz = 0;
  hipStream_t st = nullptr;
  const PassVariant* vt = nullptr;        // compiled variants of this (precision, side)
  int nvt = 0;
  int var[kPassKinds] = {0};              // variant used by each pass kind
  // (plane passes are compiled in the first variant's unit of their side)
  const PassVariant& variant_of(int kind) const {
    return is_plane_kind(kind) ? vt[0] : vt[var[tuned_kind(kind)]];
  }
  // The step plan the passes follow: the plan set, except that a masked
  // solve on the plane plan runs the ZY kinds (kPlanPlaneCut).
  int step_plan() const { return plan == GPE_PLAN_PLANE && masked ? kPlanPlaneCut : plan; }
  long long blocks(int kind, int n) const { return variant_of(kind).blocks(kind, n); }
  hipError_t launch(int kind, const void* args, int n, int batch, hipStream_t stream, hipEvent_t t0 = nullptr,
                    hipEvent_t t1 = nullptr) const {
    return variant_of(kind).launch(kind, args, n, batch, stream, t0, t1);
  }

  void* slab = nullptr;                   // psi, tables and descriptors (one allocation)
  size_t h_tlist_bytes = 0;               // bytes of each pinned time-list buffer
  void* psi = nullptr;
  void* work = nullptr;
  void* tw = nullptr;
  void* kin[3] = {nullptr, nullptr, nullptr};
  void* kin1 = nullptr;                   // separable plan: 1D factor by FFT index, 1/n folded in
  int plan = GPE_PLAN_SEPARABLE;          // step plan (GPE_PLAN_*)
  double* xa = nullptr;
  double* ka = nullptr;

  double* m_xy = nullptr;                 // FP64 static tables
  double* m_z = nullptr;
  double* m_vol[2] = {nullptr, nullptr};
  void* w_xy = nullptr;                   // working-precision static tables
  void* w_z = nullptr;
  void* w_vol[2] = {nullptr, nullptr};

  DevProgram dyn[3];                      // XY, Z, VOL
  int chunk = 0;                          // steps per dynamic-table chunk
  void* d_tab[3] = {nullptr, nullptr, nullptr};   // [chunk+1][size] slots
  void* g_tab[3] = {nullptr, nullptr, nullptr};   // diagnostic-time slot
  double* d_tlist = nullptr;
  double* h_tlist[2] = {nullptr, nullptr};
  hipEvent_t tlist_ev[2] = {nullptr, nullptr};
  int tlist_flip = 0;

  double* red = nullptr;
  long long red_cap = 0;
  unsigned int* counter = nullptr;
  double* scal = nullptr;
  double* hist = nullptr;
  long long hist_cap = 0, hist_len = 0;
  double* nkocc = nullptr;
  unsigned int* nkcnt = nullptr;
  long long nk_cap = 0;                   // blocks the nk partial buffers hold
  double* d_out = nullptr;
  double* mon = nullptr;                  // monitor images: 5 real planes + 1 complex128 plane (7 n^2 doubles)
  double* d_nkocc = nullptr;
  long long* d_nkcnt = nullptr;
  void* stage = nullptr;

  void* desc_step = nullptr;
  void* desc_diag = nullptr;
  std::vector<char> h_desc_step, h_desc_diag;

  double* red_override = nullptr;         // norm partials of the allocation-free step (the device's scratch)
  bool pending = false;                   // psi_true = psi * scal[1]
  bool work_is_spectrum = false;          // the work field holds fftn(psi) of the current psi
  bool timing = false;
  PassTimer timer;
  hipEvent_t mark[2] = {nullptr, nullptr};  // gpe_mark start / stop
  unsigned long long* stamps = nullptr;   // kMaxStamps wall-clock slots
  bool stamping = false;
  bool stamp_chained = false;             // the stream's last command was a stamped pass
  long long stamp_used = 0, stamp_last = 0, stamp_last_count = 0;
  struct StampSpan {
    int kind;
    long long begin, begin_count;         // slots whose latest value starts the span
    long long end, end_count;             // the launch's per-block end slots
  };
  std::vector<StampSpan> stamp_spans;
};

namespace {

int set_device(gpe_solve* s) {
  HIPCHK(hipSetDevice(s->ctx->device));
  return GPE_OK;
}

// ------------------------------------------------------ device block pool --
// Blocks that destroyed solves and batches held, kept per device for the next
// allocation of the same size.  A solve's buffers have sizes set by its shape,
// so sweeps and servers that create and destroy solves of a few shapes reuse
// them instead of paying one hipFree (an unmap) per buffer and one hipMalloc
// per new buffer.  Blocks enter the pool only once the device has finished
// all work on them (destroy synchronises first), leave it on an allocation of
// the same size (dzalloc then zero-fills them on the new owner's stream), and
// are released when an allocation fails, when the pool would pass its cap,
// and when a context of the device is destroyed.  gpe_mem_info counts pooled
// bytes as free.
struct BlockPool {
  std::multimap<size_t, void*> blocks;
  size_t bytes = 0;
};
std::mutex g_pool_lock;
BlockPool g_pool[kMaxDevices];
std::unordered_map<void*, size_t> g_block_size;   // bytes of every live dmalloc block
constexpr size_t kPoolCap = size_t(32) << 30;      // bytes one device's pool may hold

int current_device() {
  int d = -1;
  return hipGetDevice(&d) == hipSuccess && d >= 0 && d < kMaxDevices ? d : -1;
}

// (g_pool_lock held)
void release_pool(int dev) {
  for (auto& kv : g_pool[dev].blocks) (void)hipFree(kv.second);
  g_pool[dev].blocks.clear();
  g_pool[dev].bytes = 0;
}

size_t pooled_bytes(int dev) {
  std::lock_guard<std::mutex> guard(g_pool_lock);
  return dev >= 0 && dev < kMaxDevices ? g_pool[dev].bytes : 0;
}

int dmalloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  const int dev = current_device();
  if (dev >= 0) {
    std::lock_guard<std::mutex> guard(g_pool_lock);
    auto it = g_pool[dev].blocks.find(bytes);
    if (it != g_pool[dev].blocks.end()) {
      *p = it->second;
      g_pool[dev].blocks.erase(it);
      g_pool[dev].bytes -= bytes;
      g_block_size[*p] = bytes;
      return GPE_OK;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && dev >= 0) {
    // the pooled blocks may be what is missing
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> guard(g_pool_lock);
      release_pool(dev);
    }
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(GPE_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                   hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> guard(g_pool_lock);
  g_block_size[*p] = bytes;
  return GPE_OK;
}

// Zero-filled allocation, cleared on the owning stream `st` and complete on
// return.  (hipMemset would run on the null stream, which the non-blocking
// solve and batch streams do not wait for: the clear could land after a
// kernel on another queue had already written the buffer.)
template <typename T>
int dzalloc(T** p, size_t bytes, hipStream_t st) {
  CHK(dmalloc((void**)p, bytes));
  HIPCHK(hipMemsetAsync(*p, 0, bytes, st));
  HIPCHK(hipStreamSynchronize(st));
  return GPE_OK;
}

void dfree(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> guard(g_pool_lock);
    g_block_size.erase(p);
  }
  (void)hipFree(p);
}

// Returns a dmalloc block to device `dev`'s pool.  Only once no work on the
// device can still touch it.
void drecycle(int dev, void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> guard(g_pool_lock);
  auto it = g_block_size.find(p);
  const size_t bytes = it == g_block_size.end() ? 0 : it->second;
  if (it != g_block_size.end()) g_block_size.erase(it);
  if (bytes == 0 || dev < 0 || dev >= kMaxDevices || g_pool[dev].bytes + bytes > kPoolCap) {
    (void)hipFree(p);
    return;
  }
  g_pool[dev].blocks.emplace(bytes, p);
  g_pool[dev].bytes += bytes;
}

template <typename R>
void fill_desc(gpe_solve* s, SolveDev<R>& d, bool diag) {
  std::memset(&d, 0, sizeof(d));
  const long long nn = (long long)s->n * s->n;
  d.psi = (cpx<R>*)s->psi;
  d.work = (cpx<R>*)s->work;
  d.tw = (const cpx<double>*)s->tw;
  d.kin_x = (const cpx<double>*)s->kin[0];
  d.kin_y = (const cpx<double>*)s->kin[1];
  d.kin_z = (const cpx<double>*)s->kin[2];
  d.kin1 = (const cpx<double>*)s->kin1;
  d.karr = s->ka;
  d.xarr = s->xa;
  if (s->dyn[0].active) {
    d.pxy = (const R*)(diag ? s->g_tab[0] : s->d_tab[0]);
    d.pxy_ss = diag ? 0 : nn;
  } else {
    d.pxy = (const R*)s->w_xy;
    d.pxy_ss = 0;
  }
  if (s->dyn[1].active) {
    d.pz = (const R*)(diag ? s->g_tab[1] : s->d_tab[1]);
    d.pz_ss = diag ? 0 : s->n;
  } else {
    d.pz = (const R*)s->w_z;
    d.pz_ss = 0;
  }
  d.vre = (const R*)s->w_vol[0];
  d.vim = (const R*)s->w_vol[1];
  if (s->dyn[2].active) {
    d.vdyn = (const R*)(diag ? s->g_tab[2] : s->d_tab[2]);
    d.vdyn_ss = diag ? 0 : s->n3;
  }
  d.red = s->red;
  d.scal = s->scal;
  d.hist = s->hist;
  d.stamps = s->stamps;
  d.nkocc = s->nkocc;
  d.nkcnt = s->nkcnt;
  d.dx3 = s->dx3;
  d.kcut2 = s->kcut * s->kcut;
  d.beta = R(s->beta);
  d.hdt = R(s->dt / 2);
  d.bhdt = R(s->beta * (s->dt / 2));
  d.n = s->n;
  d.nbins = s->nbins;
  d.nview = kViewBins;
  d.vh = s->n / 2;
  d.imag = s->kind == GPE_IMAG ? 1 : 0;
  d.masked = s->masked;
}

template <typename R>
int refresh_desc_t(gpe_solve* s) {
  s->h_desc_step.resize(sizeof(SolveDev<R>));
  s->h_desc_diag.resize(sizeof(SolveDev<R>));
  fill_desc<R>(s, *reinterpret_cast<SolveDev<R>*>(s->h_desc_step.data()), false);
  fill_desc<R>(s, *reinterpret_cast<SolveDev<R>*>(s->h_desc_diag.data()), true);
  HIPCHK(hipMemcpyAsync(s->desc_step, s->h_desc_step.data(), sizeof(SolveDev<R>), hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(s->desc_diag, s->h_desc_diag.data(), sizeof(SolveDev<R>), hipMemcpyHostToDevice, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int refresh_desc(gpe_solve* s) { return s->prec == GPE_C64 ? refresh_desc_t<float>(s) : refresh_desc_t<double>(s); }

int ensure_red(gpe_solve* s, long long need) {
  if (need <= s->red_cap) return GPE_OK;
  HIPCHK(hipStreamSynchronize(s->st));
  dfree(s->red);
  s->red = nullptr;
  s->red_cap = 0;
  CHK(dzalloc(&s->red, need * sizeof(double), s->st));
  s->red_cap = need;
  return refresh_desc(s);
}

int norm_kind(int plan);

int pass(gpe_solve* s, int kind, bool diag, int slot_a, int slot_b, int flags, int hist_index = 0) {
  s->work_is_spectrum = false;            // every pass changes psi or the work field
  int stamp_index = 0;
  const long long nblk = s->stamping && kind < kPassKinds ? s->blocks(kind, s->n) : 0;
  if (s->stamping && kind < kPassKinds && s->stamp_used + 1 + nblk <= kMaxStamps) {
    // the span begins at the previous timed launch's block ends, or at a
    // reference time when other work ran since
    long long begin = s->stamp_last, begin_count = s->stamp_last_count;
    if (!s->stamp_chained) {
      begin = s->stamp_used++;
      begin_count = 1;
      launch_stamp_now(s->stamps + begin, s->st);
    }
    stamp_index = (int)s->stamp_used;
    s->stamp_used += nblk;
    s->stamp_spans.push_back({kind, begin, begin_count, stamp_index, nblk});
    s->stamp_last = stamp_index;
    s->stamp_last_count = nblk;
    // the norm finisher follows a norm pass on the stream (unless deferred)
    s->stamp_chained = !((flags & F_NORM) && !(flags & F_NORM_DEFER));
    flags |= F_STAMP;
  } else if (s->stamping) {
    // out of slots, or a kind outside the step kinds: this launch goes
    // untimed, and a later launch must not count it, so its span starts at a
    // new reference time
    s->stamp_chained = false;
  }
  flags |= s->kind == GPE_IMAG ? F_KIN1_LDS : F_KIN1_RT;
  const int parts = (flags & F_NORM_FOLD) ? int(s->blocks(norm_kind(s->step_plan()), s->n)) : 0;
  PassArgsRaw a{diag ? s->desc_diag : s->desc_step, slot_a, slot_b, flags, hist_index, stamp_index, parts,
                s->red_override};
  hipEvent_t start = nullptr, stop = nullptr;
  if (s->timing && kind < kPassKinds) HIPCHK(s->timer.begin(kind, s->st, &start, &stop));
  else if (s->timing) s->timer.chained = false;   // an untimed kind: the next span starts after it
  hipError_t e = s->launch(kind, &a, s->n, 1, s->st, start, stop);
  if (e != hipSuccess) return fail(GPE_ERR_HIP, std::string("pass launch: ") + hipGetErrorString(e));
  return GPE_OK;
}

template <typename R>
int materialize_t(gpe_solve* s) {
  if (!s->pending) return GPE_OK;
  launch_scale<R>(s->psi, s->n3, s->scal, s->st);
  HIPCHK(hipGetLastError());
  s->pending = false;
  return GPE_OK;
}
int materialize(gpe_solve* s) { return s->prec == GPE_C64 ? materialize_t<float>(s) : materialize_t<double>(s); }

long long target_size(gpe_solve* s, int target) {
  if (target == GPE_TGT_XY) return (long long)s->n * s->n;
  if (target == GPE_TGT_Z) return s->n;
  return s->n3;
}

// Checks a postfix program: opcodes, constant indices, stack discipline and
// the variables allowed for the target.
int validate_program(const int32_t* code, int32_t ncode, int32_t nconst, int target, bool allow_t) {
  if (!code || ncode <= 0 || ncode > kVmMaxCode) return fail(GPE_ERR_INVALID, "program: bad code length");
  int sp = 0;
  for (int pc = 0; pc < ncode; ++pc) {
    const int op = code[pc];
    switch (op) {
      case OP_CONST:
        if (pc + 1 >= ncode || code[pc + 1] < 0 || code[pc + 1] >= nconst)
          return fail(GPE_ERR_INVALID, "program: constant index out of range");
        ++pc;
        ++sp;
        break;
      case OP_X:
      case OP_Y:
        if (target == GPE_TGT_Z) return fail(GPE_ERR_INVALID, "program: z-table program uses x or y");
        ++sp;
        break;
      case OP_Z:
        if (target == GPE_TGT_XY) return fail(GPE_ERR_INVALID, "program: xy-table program uses z");
        ++sp;
        break;
      case OP_T:
        if (!allow_t) return fail(GPE_ERR_INVALID, "program: static program uses t");
        ++sp;
        break;
      case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_POW: case OP_MAX:
        if (sp < 2) return fail(GPE_ERR_INVALID, "program: stack underflow");
        --sp;
        break;
      case OP_NEG: case OP_SIN: case OP_COS: case OP_TAN: case OP_EXP: case OP_SQRT: case OP_TANH:
      case OP_FABS: case OP_LOG:
        if (sp < 1) return fail(GPE_ERR_INVALID, "program: stack underflow");
        break;
      default:
        return fail(GPE_ERR_INVALID, "program: unknown opcode " + std::to_string(op));
    }
    if (sp > kVmStack) return fail(GPE_ERR_INVALID, "program: expression too deep");
  }
  if (sp != 1) return fail(GPE_ERR_INVALID, "program: must leave exactly one value");
  return GPE_OK;
}

int upload_program(DevProgram& p, const int32_t* code, int32_t ncode, const double* consts, int32_t nconst) {
  dfree(p.code);
  dfree(p.consts);
  dfree(p.hv);
  p.code = nullptr;
  p.consts = nullptr;
  p.hv = nullptr;
  p.jit = nullptr;
  CHK(dmalloc((void**)&p.code, sizeof(int) * ncode));
  CHK(dmalloc((void**)&p.consts, sizeof(double) * std::max(1, nconst)));
  HIPCHK(hipMemcpy(p.code, code, sizeof(int) * ncode, hipMemcpyHostToDevice));
  if (nconst > 0) HIPCHK(hipMemcpy(p.consts, consts, sizeof(double) * nconst, hipMemcpyHostToDevice));
  p.ncode = ncode;
  p.active = true;
  return GPE_OK;
}

int ensure_volume(gpe_solve* s, int comp) {
  if (s->m_vol[comp]) return GPE_OK;
  CHK(dzalloc(&s->m_vol[comp], s->n3 * sizeof(double), s->st));
  CHK(dzalloc(&s->w_vol[comp], s->n3 * s->rsz, s->st));
  return refresh_desc(s);
}

int recast(gpe_solve* s, const double* master, void* out, long long size) {
  if (s->prec == GPE_C64)
    launch_cast<float>(master, out, size, s->st);
  else
    launch_cast<double>(master, out, size, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

// (Re)allocate the time-dependent slot tables for the active programs.
int layout_dynamic(gpe_solve* s) {
  HIPCHK(hipStreamSynchronize(s->st));
  long long biggest = 0;
  for (int g = 0; g < 3; ++g)
    if (s->dyn[g].active) biggest = std::max(biggest, target_size(s, g == 2 ? GPE_TGT_VOL : g) * (long long)s->rsz);
  int chunk = (int)std::max(1LL, std::min(256LL, kDynBudget / std::max(1LL, biggest) - 1));
  s->chunk = chunk;
  for (int g = 0; g < 3; ++g) {
    dfree(s->d_tab[g]);
    dfree(s->g_tab[g]);
    s->d_tab[g] = s->g_tab[g] = nullptr;
    if (!s->dyn[g].active) continue;
    const long long sz = target_size(s, g == 2 ? GPE_TGT_VOL : g);
    CHK(dzalloc(&s->d_tab[g], (size_t)(chunk + 1) * sz * s->rsz, s->st));
    CHK(dzalloc(&s->g_tab[g], (size_t)sz * s->rsz, s->st));
  }
  dfree(s->d_tlist);
  s->d_tlist = nullptr;
  CHK(dzalloc(&s->d_tlist, (size_t)(chunk + 2) * sizeof(double), s->st));
  for (int f = 0; f < 2; ++f) {
    // (the stream has finished: no copy from the old buffers is in flight)
    give_pinned(s->ctx, s->h_tlist_bytes, s->h_tlist[f]);
    s->h_tlist[f] = nullptr;
    HIPCHK(take_pinned(s->ctx, (size_t)(chunk + 2) * sizeof(double), (void**)&s->h_tlist[f]));
    if (!s->tlist_ev[f]) HIPCHK(hipEventCreateWithFlags(&s->tlist_ev[f], hipEventDisableTiming));
  }
  s->h_tlist_bytes = (size_t)(chunk + 2) * sizeof(double);
  return refresh_desc(s);
}

bool any_dynamic(gpe_solve* s) { return s->dyn[0].active || s->dyn[1].active || s->dyn[2].active; }

// Evaluate every active time-dependent program at times t[0..nslots) into
// slot tables (diag == false) or the single diagnostic slot (diag == true).
// Evaluates the time-dependent tables at times t[0..nslots) into the step
// slots (or the diagnostic slot), on `stream` (the solve's own by default;
// a batch uses its stream so the tables stay ordered with its passes).
int eval_dynamic(gpe_solve* s, const double* t, int nslots, bool diag, hipStream_t stream = nullptr) {
  hipStream_t st = stream ? stream : s->st;
  s->timer.chained = false;    // the next timed pass does not start at a pass's end
  s->stamp_chained = false;
  const int f = s->tlist_flip;
  s->tlist_flip ^= 1;
  HIPCHK(hipEventSynchronize(s->tlist_ev[f]));
  std::memcpy(s->h_tlist[f], t, sizeof(double) * nslots);
  double* dst_t = s->d_tlist;
  HIPCHK(hipMemcpyAsync(dst_t, s->h_tlist[f], sizeof(double) * nslots, hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(s->tlist_ev[f], st));
  for (int g = 0; g < 3; ++g) {
    if (!s->dyn[g].active) continue;
    const int target = g == 2 ? GPE_TGT_VOL : g;
    const long long size = target_size(s, target);
    const double* master = g == 0 ? s->m_xy : (g == 1 ? s->m_z : nullptr);
    void* out = diag ? s->g_tab[g] : s->d_tab[g];
    if (s->dyn[g].jit) {
      HIPCHK(jit_eval(s->dyn[g].jit, s->prec == GPE_C64, s->dyn[g].consts, target, s->n, s->xa, size, master, dst_t,
                      nslots, s->dyn[g].hv, out, st));
      continue;
    }
    if (s->prec == GPE_C64)
      launch_vm_dynamic<float>(s->dyn[g].code, s->dyn[g].ncode, s->dyn[g].consts, target, s->n, s->xa, size,
                               master, dst_t, nslots, out, st);
    else
      launch_vm_dynamic<double>(s->dyn[g].code, s->dyn[g].ncode, s->dyn[g].consts, target, s->n, s->xa, size,
                                master, dst_t, nslots, out, st);
    HIPCHK(hipGetLastError());
  }
  return GPE_OK;
}

// Pass sequence of steps [q0, q0 + count) of a call of `nsteps` steps,
// emitted through pass_fn(kind, slot_a, slot_b, flags, hist_index).  With
// time-dependent tables (`dyn`) slot q - q0 holds step q's midpoint time, so
// a chunk of `count` steps reads count + 1 slots; otherwise every slot is 0.
//
// Separable plan (GPE_PLAN_SEPARABLE), 3 sweeps per step:
//   real:      [ZK_FIRST(0)] YK  XK ZK_MID(q, q+1)  YK  XK ...  YK XK_LAST(n-1)
//   imaginary: ZK_FIRST(prescale) YK XK_LAST(norm)     per step
// Plane plan (GPE_PLAN_PLANE): the separable plan with each Z kind and the YK
// after it in one plane pass, 2 sweeps per step:
//   real:      [ZYK_FIRST(0)] XK ZYK_MID(q, q+1)  XK ...  XK_LAST(n-1)
//   imaginary: ZYK_FIRST(prescale) XK_LAST(norm)     per step
// Four-sweep plan (GPE_PLAN_FOUR_SWEEP, the momentum cutoff's 3D factor):
//   real:      [Z_FIRST(0)] Y_FWD X_KIN Y_INV Z_MID(q, q+1) ... Z_LAST(n-1)
//   imaginary: Z_FIRST(prescale) Y_FWD X_KIN Y_INV Z_LAST(norm)
// Plane plan of a masked solve (kPlanPlaneCut): the four-sweep plan with
// Y_INV, the Z kind and Y_FWD in one plane pass, 2 sweeps per real step:
//   real:      [ZY_FIRST(0)] X_KIN ZY_MID(q, q+1)  X_KIN ...  ZY_LAST(n-1)
// (the cutoff is real-time only, so imaginary steps never take this plan)
// Real time fuses the closing half-step of step q with the opening one of
// step q+1; imaginary time ends every step with the FP64 norm and defers the
// 1/sqrt(norm) scale to the next load (F_PRESCALE).
template <class P>
int emit_steps(int kind, int plan, int64_t q0, int64_t count, int64_t nsteps, bool prescale_first, bool dyn,
               bool fold, P&& pass_fn) {
  const bool sep = plan == GPE_PLAN_SEPARABLE;
  const bool plane = plan == GPE_PLAN_PLANE;
  const bool cut = plan == kPlanPlaneCut;
  for (int64_t q = q0; q < q0 + count; ++q) {
    const int sa = dyn ? (int)(q - q0) : 0, sb = dyn ? sa + 1 : 0;
    const bool first = q == 0, last = q == nsteps - 1;
    if (kind == GPE_REAL) {
      const int pre = first && prescale_first ? F_PRESCALE : 0;
      if (cut) {
        if (first) CHK(pass_fn(K_ZY_FIRST, sa, 0, pre, 0));
        CHK(pass_fn(K_X_KIN, 0, 0, 0, 0));
        if (last) CHK(pass_fn(K_ZY_LAST, sa, 0, 0, 0));
        else CHK(pass_fn(K_ZY_MID, sa, sb, 0, 0));
      } else if (plane) {
        if (first) CHK(pass_fn(K_ZYK_FIRST, sa, 0, pre, 0));
        if (last) {
          CHK(pass_fn(K_XK_LAST, sa, 0, 0, 0));
        } else {
          CHK(pass_fn(K_XK, 0, 0, 0, 0));
          CHK(pass_fn(K_ZYK_MID, sa, sb, 0, 0));
        }
      } else if (sep) {
        if (first) CHK(pass_fn(K_ZK_FIRST, sa, 0, pre, 0));
        CHK(pass_fn(K_YK, 0, 0, 0, 0));
        if (last) {
          CHK(pass_fn(K_XK_LAST, sa, 0, 0, 0));
        } else {
          CHK(pass_fn(K_XK, 0, 0, 0, 0));
          CHK(pass_fn(K_ZK_MID, sa, sb, 0, 0));
        }
      } else {
        if (first) CHK(pass_fn(K_Z_FIRST, sa, 0, pre, 0));
        CHK(pass_fn(K_Y_FWD, 0, 0, 0, 0));
        CHK(pass_fn(K_X_KIN, 0, 0, 0, 0));
        CHK(pass_fn(K_Y_INV, 0, 0, 0, 0));
        if (last) CHK(pass_fn(K_Z_LAST, sa, 0, 0, 0));
        else CHK(pass_fn(K_Z_MID, sa, sb, 0, 0));
      }
    } else {
      // With `fold`, every step but the call's last leaves its norm partials
      // for the next step's first pass to finish (one launch fewer per step);
      // the last step's finisher publishes norm and scale for the host.
      const bool folds = fold && q > q0;
      const bool defers = fold && q + 1 < q0 + count;
      const int pre = ((first ? prescale_first : true) ? F_PRESCALE : 0) | (folds ? F_NORM_FOLD : 0);
      const int hpre = folds ? (int)(q - 1) : 0;
      const int fnorm = F_NORM | (defers ? F_NORM_DEFER : 0);
      if (plane) {
        CHK(pass_fn(K_ZYK_FIRST, 0, 0, pre, hpre));
        CHK(pass_fn(K_XK_LAST, 0, 0, fnorm, (int)q));
      } else if (sep) {
        CHK(pass_fn(K_ZK_FIRST, 0, 0, pre, hpre));
        CHK(pass_fn(K_YK, 0, 0, 0, 0));
        CHK(pass_fn(K_XK_LAST, 0, 0, fnorm, (int)q));
      } else {
        CHK(pass_fn(K_Z_FIRST, 0, 0, pre, hpre));
        CHK(pass_fn(K_Y_FWD, 0, 0, 0, 0));
        CHK(pass_fn(K_X_KIN, 0, 0, 0, 0));
        CHK(pass_fn(K_Y_INV, 0, 0, 0, 0));
        CHK(pass_fn(K_Z_LAST, 0, 0, fnorm, (int)q));
      }
    }
  }
  return GPE_OK;
}

// Default step plan.  With the momentum cutoff, which is not separable: the
// plane plan (its ZY kinds) wherever a plane pass exists (half the sweeps of
// the four-sweep plan, faster alone and in batches at every such side), else
// four-sweep.  Without: the plane plan where a plane pass exists and its
// blocks are whole waves of 8-thread pencils (complex64 at 64 and 128,
// complex128 at 64: faster alone and in batches), else separable.  At 96 (4 threads a pencil,
// 24 values each, one 6-wave block per CU) the plane pass is faster alone but
// slower in batches, so solves there keep the separable plan unless set.
// (_lib.default_plan restates this rule.)
int default_plan(const gpe_solve* s) {
  if (s->masked) return s->blocks(K_ZY_MID, s->n) > 0 ? GPE_PLAN_PLANE : GPE_PLAN_FOUR_SWEEP;
  if (s->n % 64 == 0 && s->blocks(K_ZYK_MID, s->n) > 0) return GPE_PLAN_PLANE;
  return GPE_PLAN_SEPARABLE;
}

// The pass kind that writes an imaginary-time step's norm partials.
int norm_kind(int plan) { return plan == GPE_PLAN_FOUR_SWEEP ? K_Z_LAST : K_XK_LAST; }

// Whether an imaginary step's first pass takes over the norm finisher of the
// step before it (F_NORM_FOLD): every block of that pass sums all the
// partials, so only where they are few (small grids, where the finisher's
// launch is a large share of a step), and the pass's blocks must be whole
// waves.
constexpr long long kFoldParts = 1024;
bool norm_fold_ok(const gpe_solve* s) {
  if (s->blocks(norm_kind(s->step_plan()), s->n) > kFoldParts) return false;
  if (s->plan == GPE_PLAN_PLANE) return true;   // a plane block is n pencils of TP threads: whole waves
  const PassVariant& v = s->variant_of(s->plan == GPE_PLAN_SEPARABLE ? K_ZK_FIRST : K_Z_FIRST);
  return (v.tp * v.b) % 64 == 0;
}

int step_real(gpe_solve* s, int64_t nsteps, const double* tmid) {
  const bool pre = s->pending;
  s->pending = false;
  auto fn = [&](int k, int sa, int sb, int f, int h) { return pass(s, k, false, sa, sb, f, h); };
  if (!any_dynamic(s)) return emit_steps(GPE_REAL, s->step_plan(), 0, nsteps, nsteps, pre, false, false, fn);
  int64_t done = 0;
  while (done < nsteps) {
    const int64_t cs = std::min<int64_t>(s->chunk, nsteps - done);
    const int nslots = (int)std::min<int64_t>(cs + 1, nsteps - done);
    CHK(eval_dynamic(s, tmid + done, nslots, false));
    CHK(emit_steps(GPE_REAL, s->step_plan(), done, cs, nsteps, pre, true, false, fn));
    done += cs;
  }
  return GPE_OK;
}

// The reference's real-time step as separate launches, one per reference kernel
// and per transform axis (gpe/make_propagator.py:38-101): half-step, FFT z, y, x
// (cuFFT's fftn), kinetic factor, inverse FFT x, y, z (ifftn, 1/N^3 folded into
// the kinetic factor), half-step.  9 read + write sweeps of psi per step
// against 4 for the fused plan; same results to rounding.  Both half-steps of
// step i use t_mid[i], as the reference does.
int step_real_unfused(gpe_solve* s, int64_t nsteps, const double* tmid) {
  const bool dyn = any_dynamic(s);
  int64_t done = 0;
  while (done < nsteps) {
    const int64_t cs = dyn ? std::min<int64_t>(s->chunk, nsteps - done) : nsteps - done;
    if (dyn) CHK(eval_dynamic(s, tmid + done, (int)cs, false));
    for (int64_t q = 0; q < cs; ++q) {
      const int sa = dyn ? (int)q : 0;
      CHK(pass(s, K_Z_HALF, false, sa, 0, s->pending ? F_PRESCALE : 0));
      s->pending = false;
      CHK(pass(s, K_Z_FWD, false, 0, 0, 0));
      CHK(pass(s, K_Y_FWD, false, 0, 0, 0));
      CHK(pass(s, K_X_FWD, false, 0, 0, 0));
      if (s->prec == GPE_C64)
        launch_kinetic_points<float>(s->psi, s->kin[0], s->kin[1], s->kin[2], s->ka, s->n, s->masked,
                                     s->kcut * s->kcut, s->st);
      else
        launch_kinetic_points<double>(s->psi, s->kin[0], s->kin[1], s->kin[2], s->ka, s->n, s->masked,
                                      s->kcut * s->kcut, s->st);
      HIPCHK(hipGetLastError());
      s->timer.chained = false;
      s->stamp_chained = false;
      CHK(pass(s, K_X_INV, false, 0, 0, 0));
      CHK(pass(s, K_Y_INV, false, 0, 0, 0));
      CHK(pass(s, K_Z_INV, false, 0, 0, 0));
      CHK(pass(s, K_Z_HALF, false, sa, 0, 0));
    }
    done += cs;
  }
  return GPE_OK;
}

int ensure_hist(gpe_solve* s, int64_t nsteps) {
  if (nsteps <= s->hist_cap) return GPE_OK;
  HIPCHK(hipStreamSynchronize(s->st));
  dfree(s->hist);
  s->hist = nullptr;
  s->hist_cap = 0;
  const long long cap = std::max<long long>(nsteps, 1024);
  CHK(dzalloc(&s->hist, cap * sizeof(double), s->st));
  s->hist_cap = cap;
  return refresh_desc(s);
}

int step_imag(gpe_solve* s, int64_t nsteps) {
  CHK(ensure_hist(s, nsteps));
  CHK(ensure_red(s, s->blocks(norm_kind(s->step_plan()), s->n)));
  const bool pre = s->pending;
  CHK(emit_steps(GPE_IMAG, s->step_plan(), 0, nsteps, nsteps, pre, false, norm_fold_ok(s), [&](int k, int sa, int sb, int f, int h) {
    return pass(s, k, false, sa, sb, f, h);
  }));
  s->pending = true;
  s->hist_len = nsteps;
  return GPE_OK;
}

int do_step(gpe_solve* s, int64_t nsteps, const double* tmid) {
  if (nsteps < 0) return fail(GPE_ERR_INVALID, "nsteps must be >= 0");
  if (nsteps == 0) return GPE_OK;
  if (any_dynamic(s) && !tmid)
    return fail(GPE_ERR_INVALID, "t_mid is required: a time-dependent potential is registered");
  CHK(set_device(s));
  if (s->kind == GPE_REAL) return step_real(s, nsteps, tmid);
  return step_imag(s, nsteps);
}

template <typename R>
int norm_t(gpe_solve* s) {
  CHK(materialize_t<R>(s));
  CHK(ensure_red(s, kNormBlocks));
  launch_norm<R>(s->psi, s->n3, s->red, s->counter, s->scal, s->dx3, 0, s->st);
  HIPCHK(hipGetLastError());
  return GPE_OK;
}
int norm_any(gpe_solve* s) { return s->prec == GPE_C64 ? norm_t<float>(s) : norm_t<double>(s); }

int ensure_diag(gpe_solve* s) {
  bool changed = false;
  if (!s->work) {
    CHK(dzalloc(&s->work, s->n3 * s->esz, s->st));
    CHK(dzalloc(&s->d_out, 16 * sizeof(double), s->st));
    CHK(dzalloc(&s->d_nkocc, (s->nbins + kViewBins) * sizeof(double), s->st));
    CHK(dzalloc(&s->d_nkcnt, (s->nbins + kViewBins) * sizeof(long long), s->st));
    changed = true;
  }
  // partial buffers follow the blocks of the pass variants in use
  const long long bx = s->blocks(K_X_DIAG, s->n);
  if (bx > s->nk_cap) {
    HIPCHK(hipStreamSynchronize(s->st));
    dfree(s->nkocc);
    dfree(s->nkcnt);
    s->nkocc = nullptr;
    s->nkcnt = nullptr;
    s->nk_cap = 0;
    CHK(dzalloc(&s->nkocc, (size_t)bx * (s->nbins + kViewBins) * sizeof(double), s->st));
    CHK(dzalloc(&s->nkcnt, (size_t)bx * (s->nbins + kViewBins) * sizeof(unsigned int), s->st));
    s->nk_cap = bx;
    changed = true;
  }
  const long long need = std::max(s->blocks(K_Z_DIAG, s->n) * 3, bx * 2);
  if (need > s->red_cap) {
    CHK(ensure_red(s, need));
    changed = false;   // ensure_red refreshed the descriptors
  }
  return changed ? refresh_desc(s) : GPE_OK;
}

int ensure_stage(gpe_solve* s) {
  if (s->stage) return GPE_OK;
  return dmalloc(&s->stage, kStageElems * 16);
}

int check_range(gpe_solve* s, int64_t offset, int64_t count) {
  if (offset < 0 || count < 0 || offset + count > s->n3) return fail(GPE_ERR_INVALID, "element range out of bounds");
  return GPE_OK;
}

// Copies elements [offset, offset + count) of a field of the solve (psi or
// the work field) to host memory as complex128, through the staging buffer
// when the field is complex64.
int download_field(gpe_solve* s, const void* field, double* host, int64_t offset, int64_t count) {
  if (s->prec == GPE_C128) {
    HIPCHK(hipMemcpyAsync(host, (const char*)field + offset * 16, count * 16, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return GPE_OK;
  }
  CHK(ensure_stage(s));
  for (int64_t done = 0; done < count; done += kStageElems) {
    const int64_t c = std::min<int64_t>(kStageElems, count - done);
    launch_to_c128<float>((const char*)field + (offset + done) * 8, s->stage, c, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host + 2 * done, s->stage, c * 16, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  }
  return GPE_OK;
}

// F = fftn(psi), raw and unnormalised like numpy's, into the work field; psi
// keeps its values (a pending renormalisation is applied to it first).
int transform_into_work(gpe_solve* s) {
  CHK(ensure_diag(s));
  CHK(materialize(s));
  CHK(pass(s, K_Z_FWD, true, 0, 0, F_OUT_WORK));
  CHK(pass(s, K_Y_FWD, true, 0, 0, F_IN_WORK | F_OUT_WORK));
  CHK(pass(s, K_X_FWD, true, 0, 0, F_IN_WORK | F_OUT_WORK));
  s->work_is_spectrum = true;
  return GPE_OK;
}

}  // namespace

// ===================================================================== ABI ==
extern "C" {

const char* gpe_version(void) { return "gpe_hip 0.1 (gfx950)"; }

const char* gpe_last_error(void) { return g_err.c_str(); }

int gpe_set_jit(int32_t enable) {
  g_jit_enabled.store(enable ? 1 : 0);
  return GPE_OK;
}

int gpe_jit_stats(int64_t* compiles, int64_t* launches, int64_t* failures) {
  long long c = 0, l = 0, f = 0;
  jit_stats(&c, &l, &f);
  if (compiles) *compiles = c;
  if (launches) *launches = l;
  if (failures) *failures = f;
  return GPE_OK;
}

int gpe_ctx_create(int device, gpe_ctx** out) {
  if (!out) return fail(GPE_ERR_INVALID, "out is null");
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(GPE_ERR_INVALID, "no HIP device " + std::to_string(device));
  if (device >= kMaxDevices) return fail(GPE_ERR_INVALID, "device index beyond the library's table");
  HIPCHK(hipSetDevice(device));
  double* scratch_ptr = nullptr;
  {
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch_ptr[device]) {
      void* p = nullptr;
      HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p) return fail(GPE_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
      g_scratch_ptr[device] = static_cast<double*>(p);
    }
    scratch_ptr = g_scratch_ptr[device];
  }
  gpe_ctx* c = new gpe_ctx();
  c->device = device;
  c->scratch = scratch_ptr;
  *out = c;
  return GPE_OK;
}

int gpe_ctx_destroy(gpe_ctx* ctx) {
  if (!ctx) return GPE_OK;
  if (ctx->device >= 0 && ctx->device < kMaxDevices) {
    (void)hipSetDevice(ctx->device);
    std::lock_guard<std::mutex> guard(g_pool_lock);
    release_pool(ctx->device);
  }
  for (hipStream_t st : ctx->spare_streams) (void)hipStreamDestroy(st);
  for (auto& kv : ctx->spare_pinned) (void)hipHostFree(kv.second);
  delete ctx;
  return GPE_OK;
}

int gpe_mem_info(gpe_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
  if (!ctx || !free_bytes || !total_bytes) return fail(GPE_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
  *free_bytes += pooled_bytes(ctx->device);   // pooled blocks serve the next allocations
  return GPE_OK;
}

int gpe_device_info(gpe_ctx* ctx, char* name, int32_t name_len, int32_t* runtime_version,
                    int32_t* driver_version) {
  if (!ctx || !name || name_len < 1 || !runtime_version || !driver_version)
    return fail(GPE_ERR_INVALID, "null argument");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
  // marketing name and architecture; the architecture alone when the
  // runtime reports no name
  std::string text(prop.name, strnlen(prop.name, sizeof(prop.name)));
  const std::string arch(prop.gcnArchName, strnlen(prop.gcnArchName, sizeof(prop.gcnArchName)));
  text = text.empty() ? arch : text + " (" + arch + ")";
  const size_t len = std::min(text.size(), (size_t)name_len - 1);
  memcpy(name, text.data(), len);
  name[len] = '\0';
  int rt = 0, drv = 0;
  HIPCHK(hipRuntimeGetVersion(&rt));
  HIPCHK(hipDriverGetVersion(&drv));
  *runtime_version = rt;
  *driver_version = drv;
  return GPE_OK;
}

int gpe_pass_variants(int32_t n, int32_t precision, int32_t* count) {
  if (!count) return fail(GPE_ERR_INVALID, "null argument");
  if (precision != GPE_C64 && precision != GPE_C128) return fail(GPE_ERR_INVALID, "bad precision");
  const PassVariant* first = nullptr;
  *count = variants_for(precision, n, &first);
  return GPE_OK;
}

int gpe_pass_variant_info(int32_t n, int32_t precision, int32_t variant, int32_t* threads_per_pencil,
                          int32_t* pencils_per_block, int32_t* waves) {
  if (!threads_per_pencil || !pencils_per_block || !waves) return fail(GPE_ERR_INVALID, "null argument");
  if (precision != GPE_C64 && precision != GPE_C128) return fail(GPE_ERR_INVALID, "bad precision");
  const PassVariant* first = nullptr;
  const int count = variants_for(precision, n, &first);
  if (variant < 0 || variant >= count) return fail(GPE_ERR_INVALID, "variant out of range");
  *threads_per_pencil = first[variant].tp;
  *pencils_per_block = first[variant].b;
  *waves = first[variant].waves;
  return GPE_OK;
}

int gpe_solve_set_variants(gpe_solve* s, const int32_t* variants, int32_t nkinds) {
  if (!s || !variants) return fail(GPE_ERR_INVALID, "null argument");
  if (nkinds != kPassKinds) return fail(GPE_ERR_INVALID, "nkinds must be " + std::to_string(kPassKinds));
  for (int k = 0; k < kPassKinds; ++k)
    if (variants[k] < 0 || variants[k] >= s->nvt) return fail(GPE_ERR_INVALID, "variant out of range");
  CHK(set_device(s));
  HIPCHK(hipStreamSynchronize(s->st));
  std::copy(variants, variants + kPassKinds, s->var);
  // partial buffers are sized from the block counts of the variants in use
  long long need = kNormBlocks;
  for (int k = 0; k < kPassKinds; ++k) need = std::max(need, s->blocks(k, s->n) * 3);
  return ensure_red(s, need);
}

int gpe_solve_get_variants(gpe_solve* s, int32_t* variants, int32_t nkinds) {
  if (!s || !variants) return fail(GPE_ERR_INVALID, "null argument");
  if (nkinds != kPassKinds) return fail(GPE_ERR_INVALID, "nkinds must be " + std::to_string(kPassKinds));
  std::copy(s->var, s->var + kPassKinds, variants);
  return GPE_OK;
}

int gpe_solve_set_plan(gpe_solve* s, int32_t plan) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (plan != GPE_PLAN_SEPARABLE && plan != GPE_PLAN_FOUR_SWEEP && plan != GPE_PLAN_PLANE)
    return fail(GPE_ERR_INVALID, "unknown step plan");
  if (plan == GPE_PLAN_SEPARABLE && s->masked)
    return fail(GPE_ERR_INVALID,
                "the momentum cutoff is not separable: a masked solve steps with the four-sweep or the plane plan");
  if (plan == GPE_PLAN_PLANE && s->blocks(K_ZYK_MID, s->n) == 0)
    return fail(GPE_ERR_INVALID, "no plane pass at this side and precision (a plane of the field must fit the LDS)");
  CHK(set_device(s));
  HIPCHK(hipStreamSynchronize(s->st));
  s->plan = plan;
  return ensure_red(s, s->blocks(norm_kind(s->step_plan()), s->n));
}

int gpe_solve_get_plan(gpe_solve* s, int32_t* plan) {
  if (!s || !plan) return fail(GPE_ERR_INVALID, "null argument");
  *plan = s->plan;
  return GPE_OK;
}

int gpe_ctx_set_variants(gpe_ctx* ctx, int32_t n, int32_t precision, int32_t kind, const int32_t* variants,
                         int32_t nkinds) {
  if (!ctx || !variants) return fail(GPE_ERR_INVALID, "null argument");
  if (nkinds != kPassKinds) return fail(GPE_ERR_INVALID, "nkinds must be " + std::to_string(kPassKinds));
  if (precision != GPE_C64 && precision != GPE_C128) return fail(GPE_ERR_INVALID, "bad precision");
  if (kind != GPE_REAL && kind != GPE_IMAG) return fail(GPE_ERR_INVALID, "bad kind");
  const PassVariant* first = nullptr;
  const int count = variants_for(precision, n, &first);
  std::array<int, kPassKinds> v{};
  for (int k = 0; k < kPassKinds; ++k) {
    if (variants[k] < 0 || variants[k] >= count) return fail(GPE_ERR_INVALID, "variant out of range");
    v[k] = variants[k];
  }
  ctx->variants[std::make_tuple(n, precision, kind)] = v;
  return GPE_OK;
}

int gpe_solve_footprint(const gpe_solve_desc* d, size_t* resident, size_t* diag) {
  if (!d || !resident || !diag) return fail(GPE_ERR_INVALID, "null argument");
  const size_t n = (size_t)d->n, n3 = n * n * n;
  const size_t esz = d->precision == GPE_C64 ? 8 : 16, rsz = esz / 2;
  *resident = n3 * esz + n * n * (8 + rsz) + 16 * n * 8 + (4 << 20);
  const int nb = nbins_for(d->n) + kViewBins;
  *diag = n3 * esz + (n * n) * (size_t)nb * 12 + kStageElems * 16;
  return GPE_OK;
}

int gpe_solve_create(gpe_ctx* ctx, const gpe_solve_desc* d, gpe_solve** out) {
  if (!ctx || !d || !out) return fail(GPE_ERR_INVALID, "null argument");
  if (d->n < 2 || d->n > 1024) return fail(GPE_ERR_INVALID, "n must be in 2..1024");
  if (d->precision != GPE_C64 && d->precision != GPE_C128) return fail(GPE_ERR_INVALID, "bad precision");
  if (d->kind != GPE_REAL && d->kind != GPE_IMAG) return fail(GPE_ERR_INVALID, "bad kind");
  if (!(d->dt > 0) || !std::isfinite(d->dt)) return fail(GPE_ERR_INVALID, "dt must be finite and positive");
  if (!std::isfinite(d->beta)) return fail(GPE_ERR_INVALID, "beta must be finite");
  if (!d->x_array || !d->k_array) return fail(GPE_ERR_INVALID, "x_array and k_array are required");
  if (d->kcut_enable && d->kind == GPE_IMAG)
    return fail(GPE_ERR_INVALID, "the momentum cutoff is available only for real time evolution");
  HIPCHK(hipSetDevice(ctx->device));

  gpe_solve* s = new gpe_solve();
  s->ctx = ctx;
  s->n = d->n;
  s->prec = d->precision;
  s->kind = d->kind;
  s->masked = d->kcut_enable ? 1 : 0;
  s->dt = d->dt;
  s->beta = d->beta;
  s->kcut = d->kcut;
  s->coef = d->kinetic_coef == 0.0 ? 1.0 : d->kinetic_coef;
  s->n3 = (long long)d->n * d->n * d->n;
  s->esz = s->prec == GPE_C64 ? 8 : 16;
  s->rsz = s->esz / 2;
  s->nbins = nbins_for(d->n);
  const double dx = d->x_array[1] - d->x_array[0];
  s->dx3 = dx * dx * dx;
  s->nvt = variants_for(s->prec, s->n, &s->vt);
  {
    auto it = ctx->variants.find(std::make_tuple(s->n, s->prec, s->kind));
    if (it != ctx->variants.end())
      for (int k = 0; k < kPassKinds; ++k) s->var[k] = it->second[k] < s->nvt ? it->second[k] : 0;
  }

  auto bail = [&](int code) {
    gpe_solve_destroy(s);
    return code;
  };
  hipError_t e = take_stream(ctx, &s->st);
  if (e != hipSuccess) {
    delete s;
    return fail(GPE_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  int r;
  const int n = s->n;
  // The buffers every solve holds for its whole life come from one
  // zero-filled allocation: one hipMalloc and one hipFree per solve instead
  // of sixteen (each free unmaps device memory; a batch of solves freed in a
  // row spent about 2 ms apiece that way).
  {
    void** parts[] = {&s->psi, &s->tw, &s->kin[0], &s->kin[1], &s->kin[2], &s->kin1, (void**)&s->xa,
                      (void**)&s->ka, (void**)&s->m_xy, (void**)&s->m_z, &s->w_xy, &s->w_z, (void**)&s->counter,
                      (void**)&s->scal, &s->desc_step, &s->desc_diag};
    const size_t sizes[] = {size_t(s->n3) * s->esz, size_t(n) * 16, size_t(n) * 16, size_t(n) * 16, size_t(n) * 16,
                            size_t(n) * 16, n * sizeof(double), n * sizeof(double), size_t(n) * n * sizeof(double),
                            n * sizeof(double), size_t(n) * n * s->rsz, n * s->rsz, 64, 8 * sizeof(double),
                            sizeof(SolveDev<double>), sizeof(SolveDev<double>)};
    static_assert(sizeof(parts) / sizeof(parts[0]) == sizeof(sizes) / sizeof(sizes[0]), "one size per part");
    auto aligned = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t total = 0;
    for (size_t b : sizes) total += aligned(b);
    if ((r = dzalloc(&s->slab, total, s->st))) return bail(r);
    char* at = static_cast<char*>(s->slab);
    for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i) {
      *parts[i] = at;
      at += aligned(sizes[i]);
    }
  }
  if (hipMemcpy(s->xa, d->x_array, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->ka, d->k_array, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(GPE_ERR_HIP, "setup copy failed"));
  const double inv_n3 = 1.0 / double(s->n3);
  const int imag = s->kind == GPE_IMAG ? 1 : 0;
  launch_twiddle<double>(s->tw, n, s->st);
  launch_kinetic<double>(s->kin[0], s->kin[1], s->kin[2], s->kin1, s->ka, n, s->dt, s->coef, imag, inv_n3, s->st);
  s->plan = default_plan(s);
  if (hipGetLastError() != hipSuccess) return bail(fail(GPE_ERR_HIP, "table kernels failed to launch"));
  if ((r = ensure_red(s, std::max<long long>(kNormBlocks, std::max(s->blocks(K_Z_LAST, n), s->blocks(K_XK_LAST, n))))))
    return bail(r);
  // imaginary time: the per-step norm history exists from the start, so that
  // gpe_step_noalloc can record it without allocating
  if (s->kind == GPE_IMAG && (r = ensure_hist(s, 1024))) return bail(r);
  if ((r = refresh_desc(s))) return bail(r);
  *out = s;
  return GPE_OK;
}

int gpe_solve_destroy(gpe_solve* s) {
  if (!s) return GPE_OK;
  (void)hipSetDevice(s->ctx->device);
  // the buffers go to the device's block pool, so nothing on the device may
  // still use them: the solve's stream, or a batch's that holds the solve
  // (one device-wide wait, which each hipFree would otherwise make)
  (void)hipDeviceSynchronize();
  const int dev = s->ctx->device;
  void* ptrs[] = {s->slab, s->work, s->m_vol[0], s->m_vol[1], s->w_vol[0], s->w_vol[1], s->d_tab[0], s->d_tab[1],
                  s->d_tab[2], s->g_tab[0], s->g_tab[1], s->g_tab[2], s->d_tlist, s->red, s->hist, s->nkocc,
                  s->nkcnt, s->d_out, s->d_nkocc, s->d_nkcnt, s->stage, s->stamps, s->mon};
  for (void* p : ptrs) drecycle(dev, p);
  for (auto& p : s->dyn) {
    drecycle(dev, p.code);
    drecycle(dev, p.consts);
    drecycle(dev, p.hv);
  }
  for (int f = 0; f < 2; ++f) {
    give_pinned(s->ctx, s->h_tlist_bytes, s->h_tlist[f]);
    if (s->tlist_ev[f]) (void)hipEventDestroy(s->tlist_ev[f]);
  }
  s->timer.destroy();
  for (hipEvent_t e : s->mark)
    if (e) (void)hipEventDestroy(e);
  give_stream(s->ctx, s->st);
  delete s;
  return GPE_OK;
}

int gpe_psi_upload(gpe_solve* s, const double* host, int64_t offset, int64_t count) {
  if (!s || (!host && count > 0)) return fail(GPE_ERR_INVALID, "null argument");
  s->work_is_spectrum = false;
  CHK(check_range(s, offset, count));
  CHK(set_device(s));
  CHK(materialize(s));
  if (s->prec == GPE_C128) {
    HIPCHK(hipMemcpyAsync((char*)s->psi + offset * 16, host, count * 16, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return GPE_OK;
  }
  CHK(ensure_stage(s));
  for (int64_t done = 0; done < count; done += kStageElems) {
    const int64_t c = std::min<int64_t>(kStageElems, count - done);
    HIPCHK(hipMemcpyAsync(s->stage, host + 2 * done, c * 16, hipMemcpyHostToDevice, s->st));
    launch_from_c128<float>(s->stage, (char*)s->psi + (offset + done) * 8, c, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->st));
  }
  return GPE_OK;
}

int gpe_psi_download(gpe_solve* s, double* host, int64_t offset, int64_t count) {
  if (!s || (!host && count > 0)) return fail(GPE_ERR_INVALID, "null argument");
  CHK(check_range(s, offset, count));
  CHK(set_device(s));
  CHK(materialize(s));
  if (s->prec == GPE_C128) {
    HIPCHK(hipMemcpyAsync(host, (char*)s->psi + offset * 16, count * 16, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return GPE_OK;
  }
  CHK(ensure_stage(s));
  for (int64_t done = 0; done < count; done += kStageElems) {
    const int64_t c = std::min<int64_t>(kStageElems, count - done);
    launch_to_c128<float>((char*)s->psi + (offset + done) * 8, s->stage, c, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host + 2 * done, s->stage, c * 16, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  }
  return GPE_OK;
}

int gpe_psi_copy(gpe_solve* dst, gpe_solve* src) {
  if (!dst || !src) return fail(GPE_ERR_INVALID, "null argument");
  if (dst->n != src->n) return fail(GPE_ERR_INVALID, "psi copy between different grids");
  CHK(set_device(src));
  CHK(materialize(src));
  HIPCHK(hipStreamSynchronize(src->st));
  CHK(set_device(dst));
  dst->pending = false;
  dst->work_is_spectrum = false;
  if (dst->prec == src->prec) {
    HIPCHK(hipMemcpyAsync(dst->psi, src->psi, src->n3 * src->esz, hipMemcpyDeviceToDevice, dst->st));
  } else if (src->prec == GPE_C64) {
    launch_to_c128<float>(src->psi, dst->psi, src->n3, dst->st);
  } else {
    launch_from_c128<float>(src->psi, dst->psi, src->n3, dst->st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(dst->st));
  return GPE_OK;
}

int gpe_psi_gaussian(gpe_solve* s, double sigma) {
  if (!s || !(sigma > 0)) return fail(GPE_ERR_INVALID, "sigma must be positive");
  s->work_is_spectrum = false;
  CHK(set_device(s));
  s->pending = false;
  if (s->prec == GPE_C64)
    launch_gaussian<float>(s->psi, s->xa, s->n, sigma, s->st);
  else
    launch_gaussian<double>(s->psi, s->xa, s->n, sigma, s->st);
  HIPCHK(hipGetLastError());
  CHK(norm_any(s));
  s->pending = true;
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int gpe_potential_static(gpe_solve* s, int32_t target, const int32_t* code, int32_t ncode, const double* consts,
                         int32_t nconst, double scale) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (target < GPE_TGT_XY || target > GPE_TGT_VOL_IM) return fail(GPE_ERR_INVALID, "bad target");
  if (target == GPE_TGT_VOL_IM && s->kind == GPE_IMAG)
    return fail(GPE_ERR_INVALID, "imaginary time evolution requires a real static potential");
  CHK(validate_program(code, ncode, nconst, target, false));
  CHK(set_device(s));
  DevProgram p;
  CHK(upload_program(p, code, ncode, consts, nconst));
  const long long size = target_size(s, target);
  double* master;
  void* work;
  if (target == GPE_TGT_XY) {
    master = s->m_xy;
    work = s->w_xy;
  } else if (target == GPE_TGT_Z) {
    master = s->m_z;
    work = s->w_z;
  } else {
    const int comp = target == GPE_TGT_VOL ? 0 : 1;
    CHK(ensure_volume(s, comp));
    master = s->m_vol[comp];
    work = s->w_vol[comp];
  }
  launch_vm_static(p.code, p.ncode, p.consts, target, s->n, s->xa, size, scale, master, s->st);
  HIPCHK(hipGetLastError());
  int r = recast(s, master, work, size);
  dfree(p.code);
  dfree(p.consts);
  return r;
}

int gpe_potential_dynamic(gpe_solve* s, int32_t target, const int32_t* code, int32_t ncode, const double* consts,
                          int32_t nconst) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (target < GPE_TGT_XY || target > GPE_TGT_VOL) return fail(GPE_ERR_INVALID, "bad dynamic target");
  if (s->kind == GPE_IMAG)
    return fail(GPE_ERR_INVALID, "Imaginary time evolution does not accept a time dependent potential.");
  CHK(validate_program(code, ncode, nconst, target, true));
  CHK(set_device(s));
  HIPCHK(hipStreamSynchronize(s->st));
  DevProgram& p = s->dyn[target];
  CHK(upload_program(p, code, ncode, consts, nconst));
  if (g_jit_enabled.load()) {
    p.jit = jit_get(code, ncode, nconst);
    if (p.jit && jit_hoisted(p.jit) > 0) {
      if (int r = dmalloc((void**)&p.hv, sizeof(double) * kMaxSlots * jit_hoisted(p.jit))) {
        p.jit = nullptr;
        return r;
      }
    }
  }
  return layout_dynamic(s);
}

int gpe_volume_fill(gpe_solve* s, int32_t target, double value) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (target != GPE_TGT_VOL && target != GPE_TGT_VOL_IM) return fail(GPE_ERR_INVALID, "bad volume target");
  if (target == GPE_TGT_VOL_IM && s->kind == GPE_IMAG)
    return fail(GPE_ERR_INVALID, "imaginary time evolution requires a real static potential");
  CHK(set_device(s));
  const int comp = target == GPE_TGT_VOL ? 0 : 1;
  CHK(ensure_volume(s, comp));
  launch_fill_f64(s->m_vol[comp], s->n3, value, s->st);
  HIPCHK(hipGetLastError());
  return recast(s, s->m_vol[comp], s->w_vol[comp], s->n3);
}

int gpe_volume_upload(gpe_solve* s, int32_t target, const double* host, int64_t offset, int64_t count,
                      int32_t accumulate) {
  if (!s || (!host && count > 0)) return fail(GPE_ERR_INVALID, "null argument");
  if (target != GPE_TGT_VOL && target != GPE_TGT_VOL_IM) return fail(GPE_ERR_INVALID, "bad volume target");
  if (target == GPE_TGT_VOL_IM && s->kind == GPE_IMAG)
    return fail(GPE_ERR_INVALID, "imaginary time evolution requires a real static potential");
  CHK(check_range(s, offset, count));
  CHK(set_device(s));
  const int comp = target == GPE_TGT_VOL ? 0 : 1;
  CHK(ensure_volume(s, comp));
  if (!accumulate) {
    // on the solve's stream: earlier kernels may still read the table
    HIPCHK(hipMemcpyAsync(s->m_vol[comp] + offset, host, count * sizeof(double), hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  } else {
    CHK(ensure_stage(s));
    const int64_t per = kStageElems * 2;   // doubles per staging copy
    for (int64_t done = 0; done < count; done += per) {
      const int64_t c = std::min<int64_t>(per, count - done);
      HIPCHK(hipMemcpyAsync(s->stage, host + done, c * sizeof(double), hipMemcpyHostToDevice, s->st));
      launch_add_f64(s->m_vol[comp] + offset + done, (const double*)s->stage, c, s->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(s->st));
    }
  }
  return recast(s, s->m_vol[comp], s->w_vol[comp], s->n3);
}

int gpe_volume_download(gpe_solve* s, int32_t target, double* host, int64_t offset, int64_t count) {
  if (!s || (!host && count > 0)) return fail(GPE_ERR_INVALID, "null argument");
  if (target != GPE_TGT_VOL && target != GPE_TGT_VOL_IM) return fail(GPE_ERR_INVALID, "bad volume target");
  CHK(check_range(s, offset, count));
  const int comp = target == GPE_TGT_VOL ? 0 : 1;
  if (!s->m_vol[comp]) return fail(GPE_ERR_STATE, "the solve has no such static volume");
  CHK(set_device(s));
  HIPCHK(hipMemcpyAsync(host, s->m_vol[comp] + offset, count * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int gpe_step(gpe_solve* s, int64_t nsteps, const double* t_mid) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  return do_step(s, nsteps, t_mid);
}

int gpe_step_noalloc(gpe_solve* s, int64_t nsteps, const double* t_mid) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (nsteps < 0) return fail(GPE_ERR_INVALID, "nsteps must be >= 0");
  if (nsteps == 0) return GPE_OK;
  if (any_dynamic(s) && !t_mid)
    return fail(GPE_ERR_INVALID, "t_mid is required: a time-dependent potential is registered");
  if (!s->ctx->scratch) return fail(GPE_ERR_STATE, "the context holds no scratch address");
  CHK(set_device(s));
  if (s->kind == GPE_REAL) return step_real(s, nsteps, t_mid);   // real time uses no scratch
  const long long parts = s->blocks(norm_kind(s->step_plan()), s->n);
  if (parts > kScratchDoubles)
    return fail(GPE_ERR_INVALID, "the allocation-free step holds " + std::to_string(kScratchDoubles) +
                                     " norm partials; this solve needs " + std::to_string(parts));
  if (nsteps > s->hist_cap)
    return fail(GPE_ERR_INVALID, "nsteps " + std::to_string(nsteps) + " exceeds the norm history the solve holds (" +
                                     std::to_string(s->hist_cap) + " steps); the allocation-free step cannot grow it");
  const bool pre = s->pending;
  s->red_override = s->ctx->scratch;
  const int r = emit_steps(GPE_IMAG, s->step_plan(), 0, nsteps, nsteps, pre, false, norm_fold_ok(s), [&](int k, int sa, int sb, int f, int h) {
    return pass(s, k, false, sa, sb, f, h);
  });
  s->red_override = nullptr;
  CHK(r);
  s->pending = true;
  s->hist_len = nsteps;
  return GPE_OK;
}

int gpe_step_unfused(gpe_solve* s, int64_t nsteps, const double* t_mid) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (s->kind != GPE_REAL) return fail(GPE_ERR_INVALID, "the reference-structured step is real time only");
  if (nsteps < 0) return fail(GPE_ERR_INVALID, "nsteps must be >= 0");
  if (nsteps == 0) return GPE_OK;
  if (any_dynamic(s) && !t_mid)
    return fail(GPE_ERR_INVALID, "t_mid is required: a time-dependent potential is registered");
  CHK(set_device(s));
  return step_real_unfused(s, nsteps, t_mid);
}

int gpe_step_stamped(gpe_solve* s, int64_t nsteps, const double* t_mid, double* us_per_kind,
                     int64_t* launches_per_kind) {
  if (!s || !us_per_kind || !launches_per_kind) return fail(GPE_ERR_INVALID, "null argument");
  CHK(set_device(s));
  int rate_khz = 0;
  HIPCHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, s->ctx->device));
  if (rate_khz <= 0) return fail(GPE_ERR_HIP, "the device reports no wall-clock rate");
  if (!s->stamps) {
    CHK(dmalloc((void**)&s->stamps, kMaxStamps * sizeof(unsigned long long)));
    CHK(refresh_desc(s));
  }
  launch_clear_stamps(s->stamps, kMaxStamps, s->st);
  HIPCHK(hipGetLastError());
  s->stamp_spans.clear();
  s->stamp_used = 0;
  s->stamp_chained = false;
  s->stamping = true;
  const int r = do_step(s, nsteps, t_mid);
  s->stamping = false;
  s->stamp_chained = false;
  CHK(r);
  HIPCHK(hipGetLastError());
  std::vector<unsigned long long> h(s->stamp_used);
  if (!h.empty())
    HIPCHK(hipMemcpyAsync(h.data(), s->stamps, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  for (int k = 0; k < kPassKinds; ++k) {
    us_per_kind[k] = 0.0;
    launches_per_kind[k] = 0;
  }
  auto latest = [&](long long first, long long count) {
    unsigned long long t = 0;
    for (long long i = first; i < first + count; ++i) t = std::max(t, h[i]);
    return t;
  };
  for (const auto& span : s->stamp_spans) {
    const unsigned long long begin = latest(span.begin, span.begin_count), end = latest(span.end, span.end_count);
    if (end < begin) return fail(GPE_ERR_STATE, "a stamped launch ended before its span began");
    us_per_kind[span.kind] += double(end - begin) * 1e3 / rate_khz;
    launches_per_kind[span.kind] += 1;
  }
  return GPE_OK;
}

int gpe_step_timed(gpe_solve* s, int64_t nsteps, const double* t_mid, double* ms_per_kind,
                   int64_t* launches_per_kind) {
  if (!s || !ms_per_kind || !launches_per_kind) return fail(GPE_ERR_INVALID, "null argument");
  CHK(set_device(s));
  s->timer.reset();
  s->timing = true;
  const int r = do_step(s, nsteps, t_mid);
  s->timing = false;
  const hipError_t e = s->timer.collect(s->st);
  CHK(r);
  if (e != hipSuccess) return fail(GPE_ERR_HIP, std::string("pass timing: ") + hipGetErrorString(e));
  for (int k = 0; k < kPassKinds; ++k) {
    ms_per_kind[k] = s->timer.ms[k];
    launches_per_kind[k] = s->timer.cnt[k];
  }
  return GPE_OK;
}

int gpe_step_kernel_times(gpe_solve* s, int64_t nsteps, const double* t_mid, double* us_per_kind,
                          int64_t* launches_per_kind) {
  if (!s || !us_per_kind || !launches_per_kind) return fail(GPE_ERR_INVALID, "null argument");
  CHK(set_device(s));
  s->timer.reset();
  s->timer.own = true;
  s->timing = true;
  const int r = do_step(s, nsteps, t_mid);
  s->timing = false;
  const hipError_t e = s->timer.collect(s->st);
  s->timer.own = false;
  CHK(r);
  if (e != hipSuccess) return fail(GPE_ERR_HIP, std::string("kernel timing: ") + hipGetErrorString(e));
  for (int k = 0; k < kPassKinds; ++k) {
    us_per_kind[k] = s->timer.ms[k] * 1e3;
    launches_per_kind[k] = s->timer.cnt[k];
  }
  return GPE_OK;
}

int gpe_norm_history(gpe_solve* s, double* out, int64_t count) {
  if (!s || (!out && count > 0)) return fail(GPE_ERR_INVALID, "null argument");
  if (count > s->hist_len) return fail(GPE_ERR_STATE, "fewer steps recorded than requested");
  CHK(set_device(s));
  HIPCHK(hipStreamSynchronize(s->st));
  HIPCHK(hipMemcpy(out, s->hist + (s->hist_len - count), count * sizeof(double), hipMemcpyDeviceToHost));
  return GPE_OK;
}

int gpe_norm(gpe_solve* s, double* out) {
  if (!s || !out) return fail(GPE_ERR_INVALID, "null argument");
  CHK(set_device(s));
  CHK(norm_any(s));
  HIPCHK(hipMemcpyAsync(out, s->scal, sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int gpe_normalise(gpe_solve* s) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  s->work_is_spectrum = false;
  CHK(set_device(s));
  CHK(norm_any(s));
  s->pending = true;
  return GPE_OK;
}

int gpe_diagnostics(gpe_solve* s, double t, double* scalars, int64_t* counts, double* occupation, int32_t nbins) {
  if (!s || !scalars) return fail(GPE_ERR_INVALID, "null argument");
  if ((counts || occupation) && nbins != s->nbins)
    return fail(GPE_ERR_INVALID, "nbins must be " + std::to_string(s->nbins));
  CHK(set_device(s));
  CHK(ensure_diag(s));
  if (any_dynamic(s)) CHK(eval_dynamic(s, &t, 1, true));
  const int pre = s->pending ? F_PRESCALE : 0;
  const long long bz = s->blocks(K_Z_DIAG, s->n), bx = s->blocks(K_X_DIAG, s->n);
  CHK(pass(s, K_Z_DIAG, true, 0, 0, pre | F_OUT_WORK));
  launch_finish_sums(s->red, bz, 3, s->d_out, s->st);
  HIPCHK(hipGetLastError());
  CHK(pass(s, K_Y_FWD, true, 0, 0, F_IN_WORK | F_OUT_WORK));
  const bool spectrum = counts || occupation;
  CHK(pass(s, K_X_DIAG, true, 0, 0, F_IN_WORK | (spectrum ? F_SPECTRUM : 0)));
  launch_finish_sums(s->red, bx, 2, s->d_out + 3, s->st);
  if (spectrum) launch_finish_nk(s->nkocc, s->nkcnt, bx, s->nbins + kViewBins, s->d_nkocc, s->d_nkcnt, s->st);
  HIPCHK(hipGetLastError());
  double h[5];
  HIPCHK(hipMemcpyAsync(h, s->d_out, 5 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (occupation)
    HIPCHK(hipMemcpyAsync(occupation, s->d_nkocc, s->nbins * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (counts) HIPCHK(hipMemcpyAsync(counts, s->d_nkcnt, s->nbins * sizeof(long long), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  if (spectrum) s->view_ready = true;
  const double n3 = double(s->n3);
  const double norm = s->dx3 * h[0];
  const double ekin = h[3] * s->dx3 / n3;
  const double epot = s->dx3 * h[1];
  const double eint = 0.5 * s->beta * s->dx3 * h[2];
  scalars[0] = norm;
  scalars[1] = ekin;
  scalars[2] = epot;
  scalars[3] = eint;
  scalars[4] = ekin + epot + eint;
  scalars[5] = (ekin + epot + 2.0 * eint) / norm;
  scalars[6] = h[4] * s->dx3 / n3;
  return GPE_OK;
}

int gpe_angular_momentum(gpe_solve* s, double* out) {
  if (!s || !out) return fail(GPE_ERR_INVALID, "null argument");
  CHK(set_device(s));
  const long long bx = s->blocks(K_X_LZ, s->n), by = s->blocks(K_Y_LZ, s->n);
  CHK(ensure_red(s, std::max(bx, by)));
  if (!s->d_out) CHK(dzalloc(&s->d_out, 16 * sizeof(double), s->st));
  const int pre = s->pending ? F_PRESCALE : 0;
  CHK(pass(s, K_Y_LZ, true, 0, 0, pre));
  launch_finish_sums(s->red, by, 1, s->d_out + 8, s->st);
  CHK(pass(s, K_X_LZ, true, 0, 0, pre));
  launch_finish_sums(s->red, bx, 1, s->d_out + 9, s->st);
  HIPCHK(hipGetLastError());
  double h[2];
  HIPCHK(hipMemcpyAsync(h, s->d_out + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  *out = (h[0] - h[1]) * s->dx3 / double(s->n);
  return GPE_OK;
}

// ------------------------------------------------------------- batches --
struct gpe_batch {
  gpe_ctx* ctx = nullptr;
  std::vector<gpe_solve*> solves;
  int n = 0, prec = 0, kind = 0;
  hipStream_t st = nullptr;
  void* desc = nullptr;                 // device array of per-solve descriptors
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t graph_steps = 0;
  int graph_var[kPassKinds] = {0};      // pass variants the graph was captured with
  int graph_plan = -1;                  // step plan the graph was captured with
  std::vector<char> uploaded;           // descriptor bytes last copied to `desc`
  bool timing = false;
  PassTimer timer;
};

int gpe_batch_destroy(gpe_batch* b) {
  if (!b) return GPE_OK;
  (void)hipSetDevice(b->ctx->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  drecycle(b->ctx->device, b->desc);   // (the batch's stream, the only user, has finished)
  b->timer.destroy();
  give_stream(b->ctx, b->st);
  delete b;
  return GPE_OK;
}

int gpe_batch_create(gpe_ctx* ctx, gpe_solve* const* solves, int32_t count, gpe_batch** out) {
  if (!ctx || !solves || count <= 0 || !out) return fail(GPE_ERR_INVALID, "bad batch arguments");
  for (int i = 0; i < count; ++i) {
    gpe_solve* s = solves[i];
    if (!s || s->ctx != ctx) return fail(GPE_ERR_INVALID, "batch solve missing or on another context");
    if (s->n != solves[0]->n || s->prec != solves[0]->prec || s->kind != solves[0]->kind)
      return fail(GPE_ERR_INVALID, "a batch needs solves of one side, precision and kind");
    for (int j = 0; j < i; ++j)
      if (solves[j] == s) return fail(GPE_ERR_INVALID, "a solve appears twice in the batch");
  }
  HIPCHK(hipSetDevice(ctx->device));
  gpe_batch* b = new gpe_batch();
  b->ctx = ctx;
  b->solves.assign(solves, solves + count);
  b->n = solves[0]->n;
  b->prec = solves[0]->prec;
  b->kind = solves[0]->kind;
  auto bail = [&](int code) {
    gpe_batch_destroy(b);
    return code;
  };
  if (take_stream(ctx, &b->st) != hipSuccess) return bail(fail(GPE_ERR_HIP, "batch stream creation failed"));
  const size_t dsz = b->prec == GPE_C64 ? sizeof(SolveDev<float>) : sizeof(SolveDev<double>);
  if (int r = dmalloc(&b->desc, dsz * count)) return bail(r);
  *out = b;
  return GPE_OK;
}

namespace {

// Bring every member to the state a batched step sequence expects and copy
// the descriptors into the batch's device array.
int batch_prepare(gpe_batch* b, int64_t nsteps, bool* prescale_first) {
  const size_t dsz = b->prec == GPE_C64 ? sizeof(SolveDev<float>) : sizeof(SolveDev<double>);
  *prescale_first = false;
  for (gpe_solve* s : b->solves)
    if (!std::equal(s->var, s->var + kPassKinds, b->solves[0]->var))
      return fail(GPE_ERR_INVALID, "batch members must use the same pass variants");
  for (gpe_solve* s : b->solves)
    if (s->step_plan() != b->solves[0]->step_plan())
      return fail(GPE_ERR_INVALID, "batch members must use the same step plan (and, on the plane plan, all or none "
                                   "with the momentum cutoff)");
  for (gpe_solve* s : b->solves) {
    s->work_is_spectrum = false;          // the batch advances psi
    if (b->kind == GPE_IMAG) {
      CHK(ensure_hist(s, nsteps));
      CHK(ensure_red(s, s->blocks(norm_kind(s->step_plan()), s->n)));
      if (!s->pending) {
        // a unit pending scale lets every member take the prescaled path; the
        // source is a static constant, so it outlives the asynchronous copy
        HIPCHK(hipMemcpyAsync(s->scal + 1, &kUnitScale, sizeof(double), hipMemcpyHostToDevice, s->st));
        s->pending = true;
      }
      *prescale_first = true;
    } else {
      CHK(materialize(s));
    }
    HIPCHK(hipStreamSynchronize(s->st));
  }
  std::vector<char> host(dsz * b->solves.size());
  for (size_t i = 0; i < b->solves.size(); ++i) std::memcpy(host.data() + i * dsz, b->solves[i]->h_desc_step.data(), dsz);
  if (host != b->uploaded) {   // descriptors change only when a member reallocates
    HIPCHK(hipMemcpyAsync(b->desc, host.data(), host.size(), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    b->uploaded.swap(host);
  }
  return GPE_OK;
}

int batch_pass(gpe_batch* b, int kind, int sa, int sb, int flags, int hist) {
  flags |= b->kind == GPE_IMAG ? F_KIN1_LDS : F_KIN1_RT;
  const int parts = (flags & F_NORM_FOLD) ? int(b->solves[0]->blocks(norm_kind(b->solves[0]->step_plan()), b->n)) : 0;
  PassArgsRaw a{b->desc, sa, sb, flags, hist, 0, parts, nullptr};
  hipEvent_t start = nullptr, stop = nullptr;
  if (b->timing && kind < kPassKinds) HIPCHK(b->timer.begin(kind, b->st, &start, &stop));
  else if (b->timing) b->timer.chained = false;   // an untimed kind: the next span starts after it
  hipError_t e = b->solves[0]->launch(kind, &a, b->n, (int)b->solves.size(), b->st, start, stop);
  if (e != hipSuccess) return fail(GPE_ERR_HIP, std::string("batched pass launch: ") + hipGetErrorString(e));
  return GPE_OK;
}

void batch_finish(gpe_batch* b, int64_t nsteps) {
  for (gpe_solve* s : b->solves) {
    if (b->kind == GPE_IMAG) {
      s->pending = true;
      s->hist_len = nsteps;
    } else {
      s->pending = false;
    }
  }
}

}  // namespace

int gpe_batch_launch(gpe_batch* b, int64_t nsteps, int32_t use_graph) {
  if (!b) return fail(GPE_ERR_INVALID, "null batch");
  if (nsteps < 0) return fail(GPE_ERR_INVALID, "nsteps must be >= 0");
  if (nsteps == 0) return GPE_OK;
  for (gpe_solve* s : b->solves)
    if (any_dynamic(s))
      return fail(GPE_ERR_INVALID, "a batch member carries a time-dependent program: step with gpe_batch_step_times");
  HIPCHK(hipSetDevice(b->ctx->device));
  bool pre = false;
  CHK(batch_prepare(b, nsteps, &pre));
  auto emit = [&]() {
    return emit_steps(b->kind, b->solves[0]->step_plan(), 0, nsteps, nsteps, pre, false,
                      b->kind == GPE_IMAG && norm_fold_ok(b->solves[0]), [&](int k, int sa, int sb, int f, int h) {
      return batch_pass(b, k, sa, sb, f, h);
    });
  };
  if (!use_graph || b->timing) {
    CHK(emit());
  } else {
    const int* var = b->solves[0]->var;
    if (!b->exec || b->graph_steps != nsteps || !std::equal(var, var + kPassKinds, b->graph_var) ||
        b->graph_plan != b->solves[0]->plan) {
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      const int r = emit();
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->st, &g);
      if (r != GPE_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
      }
      if (e != hipSuccess) return fail(GPE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      b->graph = g;
      HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_steps = nsteps;
      b->graph_plan = b->solves[0]->plan;
      std::copy(var, var + kPassKinds, b->graph_var);
    }
    HIPCHK(hipGraphLaunch(b->exec, b->st));
  }
  batch_finish(b, nsteps);
  return GPE_OK;
}

int gpe_batch_wait(gpe_batch* b) {
  if (!b) return fail(GPE_ERR_INVALID, "null batch");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize(b->st));
  return GPE_OK;
}

int gpe_batch_done(gpe_batch* b, int32_t* done) {
  if (!b || !done) return fail(GPE_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(b->ctx->device));
  const hipError_t e = hipStreamQuery(b->st);
  if (e == hipErrorNotReady) {
    // not an error: keep it from reading as one at the next launch check
    (void)hipGetLastError();
    *done = 0;
    return GPE_OK;
  }
  HIPCHK(e);
  *done = 1;
  return GPE_OK;
}

int gpe_batch_norm_history(gpe_batch* b, double* out, int64_t count) {
  if (!b || (!out && count > 0)) return fail(GPE_ERR_INVALID, "null argument");
  for (gpe_solve* s : b->solves)
    if (count > s->hist_len) return fail(GPE_ERR_STATE, "fewer steps recorded than requested");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize(b->st));
  for (size_t i = 0; i < b->solves.size(); ++i) {
    gpe_solve* s = b->solves[i];
    HIPCHK(hipMemcpyAsync(out + i * count, s->hist + (s->hist_len - count), count * sizeof(double),
                          hipMemcpyDeviceToHost, b->st));
  }
  HIPCHK(hipStreamSynchronize(b->st));
  return GPE_OK;
}

int gpe_batch_step(gpe_batch* b, int64_t nsteps, int32_t use_graph) {
  CHK(gpe_batch_launch(b, nsteps, use_graph));
  return gpe_batch_wait(b);
}

int gpe_batch_step_times(gpe_batch* b, int64_t nsteps, const double* t_mid, int64_t stride) {
  if (!b) return fail(GPE_ERR_INVALID, "null batch");
  if (nsteps < 0 || stride < 0) return fail(GPE_ERR_INVALID, "nsteps and stride must be >= 0");
  if (nsteps == 0) return GPE_OK;
  bool dyn = false;
  int chunk = 1 << 30;
  for (gpe_solve* s : b->solves)
    if (any_dynamic(s)) {
      dyn = true;
      chunk = std::min(chunk, s->chunk);
    }
  if (!dyn) return gpe_batch_step(b, nsteps, 1);
  if (!t_mid) return fail(GPE_ERR_INVALID, "t_mid is required when a member carries a time-dependent program");
  HIPCHK(hipSetDevice(b->ctx->device));
  bool pre = false;
  CHK(batch_prepare(b, nsteps, &pre));
  // Same pass sequence as a single real-time solve (step_real); before each
  // chunk every time-dependent member evaluates its own slots on the batch
  // stream, so the tables are ordered with the passes that read them.
  auto fn = [&](int k, int sa, int sb, int f, int h) { return batch_pass(b, k, sa, sb, f, h); };
  int64_t done = 0;
  while (done < nsteps) {
    const int64_t cs = std::min<int64_t>(chunk, nsteps - done);
    const int nslots = (int)std::min<int64_t>(cs + 1, nsteps - done);
    for (size_t i = 0; i < b->solves.size(); ++i)
      if (any_dynamic(b->solves[i])) CHK(eval_dynamic(b->solves[i], t_mid + i * stride + done, nslots, false, b->st));
    CHK(emit_steps(GPE_REAL, b->solves[0]->step_plan(), done, cs, nsteps, pre, true, false, fn));
    done += cs;
  }
  batch_finish(b, nsteps);
  HIPCHK(hipStreamSynchronize(b->st));
  return GPE_OK;
}

int gpe_batch_step_timed(gpe_batch* b, int64_t nsteps, double* ms_per_kind, int64_t* launches_per_kind) {
  if (!b || !ms_per_kind || !launches_per_kind) return fail(GPE_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(b->ctx->device));
  b->timer.reset();
  b->timing = true;
  const int r = gpe_batch_step(b, nsteps, 0);
  b->timing = false;
  const hipError_t e = b->timer.collect(b->st);
  CHK(r);
  if (e != hipSuccess) return fail(GPE_ERR_HIP, std::string("pass timing: ") + hipGetErrorString(e));
  for (int k = 0; k < kPassKinds; ++k) {
    ms_per_kind[k] = b->timer.ms[k];
    launches_per_kind[k] = b->timer.cnt[k];
  }
  return GPE_OK;
}

int gpe_sync(gpe_solve* s) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  CHK(set_device(s));
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int gpe_transform(gpe_solve* s) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  CHK(set_device(s));
  CHK(transform_into_work(s));
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int gpe_transform_download(gpe_solve* s, double* host, int64_t offset, int64_t count) {
  if (!s || (!host && count > 0)) return fail(GPE_ERR_INVALID, "null argument");
  if (!s->work_is_spectrum) return fail(GPE_ERR_STATE, "no gpe_transform of the current field");
  CHK(check_range(s, offset, count));
  CHK(set_device(s));
  return download_field(s, s->work, host, offset, count);
}

int gpe_monitor_images(gpe_solve* s, int32_t which, double* tof_xy, double* tof_yz, double* tof_slice_xy,
                       double* tof_slice_yz, double* density_xy, double* plane_mid) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (which & ~GPE_MON_ALL) return fail(GPE_ERR_INVALID, "unknown monitor image bit");
  double* outs[6] = {tof_xy, tof_yz, tof_slice_xy, tof_slice_yz, density_xy, plane_mid};
  for (int b = 0; b < 6; ++b)
    if ((which >> b & 1) && !outs[b]) return fail(GPE_ERR_INVALID, "null output for a requested image");
  if (!which) return GPE_OK;
  CHK(set_device(s));
  const long long plane = (long long)s->n * s->n;
  if (!s->mon) CHK(dmalloc((void**)&s->mon, (size_t)plane * 7 * sizeof(double)));
  if (which & (GPE_MON_TOF_XY | GPE_MON_TOF_YZ | GPE_MON_TOF_SLICE_XY | GPE_MON_TOF_SLICE_YZ)) {
    if (!s->work_is_spectrum) CHK(transform_into_work(s));
  } else {
    CHK(materialize(s));
  }
  const bool c64 = s->prec == GPE_C64;
  auto sum = [&](const void* f, int axis, double* out) {
    if (c64) launch_abs2_sum<float>(f, s->n, axis, out, s->st);
    else launch_abs2_sum<double>(f, s->n, axis, out, s->st);
  };
  auto cut = [&](const void* f, int axis, int index, void* out_c128, double* out_abs2) {
    if (c64) launch_plane<float>(f, s->n, axis, index, out_c128, out_abs2, s->st);
    else launch_plane<double>(f, s->n, axis, index, out_c128, out_abs2, s->st);
  };
  if (which & GPE_MON_TOF_XY) sum(s->work, 2, s->mon);
  if (which & GPE_MON_TOF_YZ) sum(s->work, 0, s->mon + plane);
  if (which & GPE_MON_TOF_SLICE_XY) cut(s->work, 2, 0, nullptr, s->mon + 2 * plane);
  if (which & GPE_MON_TOF_SLICE_YZ) cut(s->work, 0, 0, nullptr, s->mon + 3 * plane);
  if (which & GPE_MON_DENSITY_XY) sum(s->psi, 2, s->mon + 4 * plane);
  if (which & GPE_MON_PLANE_MID) cut(s->psi, 2, s->n / 2, s->mon + 5 * plane, nullptr);
  HIPCHK(hipGetLastError());
  for (int b = 0; b < 6; ++b) {
    if (!(which >> b & 1)) continue;
    const size_t bytes = (size_t)plane * sizeof(double) * (b == 5 ? 2 : 1);
    HIPCHK(hipMemcpyAsync(outs[b], s->mon + b * plane, bytes, hipMemcpyDeviceToHost, s->st));
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int gpe_spectrum_view(gpe_solve* s, int64_t* counts, double* occupation, int32_t nbins) {
  if (!s || (!counts && !occupation)) return fail(GPE_ERR_INVALID, "null argument");
  if (nbins != kViewBins) return fail(GPE_ERR_INVALID, "nbins must be " + std::to_string(kViewBins));
  if (!s->view_ready) return fail(GPE_ERR_STATE, "no gpe_diagnostics call has filled the view yet");
  CHK(set_device(s));
  if (occupation)
    HIPCHK(hipMemcpyAsync(occupation, s->d_nkocc + s->nbins, kViewBins * sizeof(double), hipMemcpyDeviceToHost,
                          s->st));
  if (counts)
    HIPCHK(hipMemcpyAsync(counts, s->d_nkcnt + s->nbins, kViewBins * sizeof(long long), hipMemcpyDeviceToHost,
                          s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return GPE_OK;
}

int gpe_mark(gpe_solve* s, int32_t which) {
  if (!s) return fail(GPE_ERR_INVALID, "null solve");
  if (which != 0 && which != 1) return fail(GPE_ERR_INVALID, "mark must be 0 (start) or 1 (stop)");
  CHK(set_device(s));
  if (!s->mark[which]) HIPCHK(hipEventCreate(&s->mark[which]));
  HIPCHK(hipEventRecord(s->mark[which], s->st));
  s->timer.chained = false;               // a pass timed next starts after the mark
  s->stamp_chained = false;
  return GPE_OK;
}

int gpe_mark_elapsed(gpe_solve* s, double* ms) {
  if (!s || !ms) return fail(GPE_ERR_INVALID, "null argument");
  if (!s->mark[0] || !s->mark[1]) return fail(GPE_ERR_INVALID, "both marks must be recorded first");
  CHK(set_device(s));
  HIPCHK(hipEventSynchronize(s->mark[1]));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, s->mark[0], s->mark[1]));
  *ms = t;
  return GPE_OK;
}

int gpe_bench_copy(gpe_ctx* ctx, int64_t bytes, int32_t reps, double* ms_per_copy) {
  if (!ctx || !ms_per_copy || bytes <= 0 || bytes % 16 || reps <= 0) return fail(GPE_ERR_INVALID, "bad argument");
  HIPCHK(hipSetDevice(ctx->device));
  void *a = nullptr, *b = nullptr;
  CHK(dmalloc(&a, bytes));
  if (int r = dmalloc(&b, bytes)) {
    dfree(a);
    return r;
  }
  hipStream_t st;
  hipEvent_t e0, e1;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipMemsetAsync(a, 0, bytes, st));
  launch_copy(a, b, bytes, st);
  HIPCHK(hipEventRecord(e0, st));
  for (int i = 0; i < reps; ++i) launch_copy(a, b, bytes, st);
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  dfree(a);
  dfree(b);
  return GPE_OK;
}

}  // extern "C"
