// C ABI of the MI355X ion-flight engine (include/ion_gym_amd.h): engine,
// field packs, device-resident ensembles and the flight launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ig_births.h"
#include "ig_common.h"
#include "ig_compose_kernel.h"
#include "ig_metal.h"
#include "ig_solve.h"

using namespace ig;

namespace {
int launch_route(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch,
                        int block, void *stream) {
  if (n_launch <= 0) return 0;
  switch (f.route) {
  case IG_ROUTE_RZ:
    switch (f.prec) {
    case IG_PREC_F64: return launch_rz_prec<IG_PREC_F64>(f, p, e, n_launch, block, stream);
    case IG_PREC_F32_STORE:
      return launch_rz_prec<IG_PREC_F32_STORE>(f, p, e, n_launch, block, stream);
    case IG_PREC_F32: return launch_rz_prec<IG_PREC_F32>(f, p, e, n_launch, block, stream);
    default: return -2;
    }
  case IG_ROUTE_XYZ:
    switch (f.prec) {
    case IG_PREC_F64: return launch_xyz_prec<IG_PREC_F64>(f, p, e, n_launch, block, stream);
    case IG_PREC_F32_STORE:
      return launch_xyz_prec<IG_PREC_F32_STORE>(f, p, e, n_launch, block, stream);
    case IG_PREC_F32: return launch_xyz_prec<IG_PREC_F32>(f, p, e, n_launch, block, stream);
    default: return -2;
    }
  case IG_ROUTE_PLANAR:
    switch (f.prec) {
    case IG_PREC_F64: return launch_planar_prec<IG_PREC_F64>(f, p, e, n_launch, block, stream);
    case IG_PREC_F32_STORE:
      return launch_planar_prec<IG_PREC_F32_STORE>(f, p, e, n_launch, block, stream);
    case IG_PREC_F32: return launch_planar_prec<IG_PREC_F32>(f, p, e, n_launch, block, stream);
    default: return -2;
    }
  default: return -2;
  }
}
} // namespace

namespace {

thread_local std::string g_err;

// parked hard-sphere lanes a wave scatters together when the caller leaves
// coll_batch at 0 (the autotuner measures the best value per workload);
// measured on the funnel at 2 waves/SIMD: 193 ms per window at 16, 184 at
// 24, 187 at 32, 220 at 48
constexpr int kDefaultCollBatch = 24;
// FP32-basis hard sphere: parked lanes a wave tries partners for together
// (measured on MI355X: the funnel r-z flight runs 17 % faster at 8 than at
// 1; the 11-set SLIM 3-D flight, whose step is longer, is best at 1)
constexpr int kDefaultCollBatchFlat = 8;
constexpr int kDefaultCollBatchFlat3D = 1;
// default register-budget variant (waves per SIMD) of the flight kernels per
// precision, route and gas, as measured on MI355X (1M-ion windows):
// * r-z and planar, FP32 basis: 3 (funnel HS 35.7 ms against 48.7 at 4;
//   vacuum 27.8 against 34.4 at 1; SDS drift tube 14.8 against 17.3 at
//   1), 4 for the planar vacuum step (27.5 against 30.6 at 3);
// * r-z and planar, FP64 and FP32 storage: hard sphere 2 (funnel 184 ms
//   against 308 at 1 and 327 at 3), vacuum 3 (41.0 against 48.3 at 1),
//   SDS 1 (38.7 against 71.2 at 3);
// * 3-D: the FP32 kernel, whose step keeps its cell's corner sums, at 2
//   with gas on the SLIM board (3 spills inside the step) and at the
//   compiler's own budget in vacuum (222 VGPRs whatever the budget:
//   quadrupole 6.8e9 at 1 against 5.5e9 at 2); FP64 and FP32 storage
//   with hard sphere at 2 (quadrupole 67.7 ms against 72.2 at 1 and 97.5
//   at 3), in vacuum at the compiler's own budget (67.9 against 85.3 at 3).
// The planar hard-sphere kernels were measured too (FP32 42.4 ms at 3
// against 52.7 at 1; FP64 386 at 2 against 620 at 1 and 436 at 3); the
// planar SDS kernels follow the r-z ones.
inline int default_occupancy(int prec, int route, int gas) {
  if (route == IG_ROUTE_XYZ) {
    if (prec != IG_PREC_F32) return gas == IG_GAS_HS ? 2 : 1;
    return gas == IG_GAS_VACUUM ? 1 : 2;
  }
  if (prec == IG_PREC_F32)
    return route == IG_ROUTE_PLANAR && gas == IG_GAS_VACUUM ? 4 : 3;
  if (gas == IG_GAS_HS) return 2;
  return gas == IG_GAS_VACUUM ? 3 : 1;
}

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPX(call)                                                                       \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail(IG_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e));          \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

int dev_alloc(DevBuf &b, size_t bytes) {
  b.bytes = bytes;
  if (bytes == 0) return 0;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(e == hipErrorOutOfMemory ? IG_E_NOMEM : IG_E_HIP,
                std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  return 0;
}

// ---- linked drive channels (IG_FOLD_LINKED) -----------------------------
// Given channel k weighs its grid E_k by a_k w(om_k t + ph_k) + b_k (3-D;
// the r-z and planar routes fly unit waveforms, a_k = 1 and b_k = 0, the
// amplitude sits in the grid).  Channels of one kind (sin, cos, or square
// at duty 0.5) and one frequency whose phases lie a whole number n_k of
// half periods from the first such channel's (the group's leader L) obey
// w(x + n pi) = (-1)^n w(x): the group is stored as one set
// sum_k (-1)^n_k (a_k / a_L) E_k flown with the leader's waveform and
// amplitude, and the groups' offsets as one constant set sum_k b_k E_k of
// weight 1.  Other channels are stored as given.
struct FoldPlan {
  bool on = false;
  bool affine = false;          // 3-D weights a w + b, else unit waveforms
  int K0 = 0, K1 = 0;           // drive channels given, stored
  std::vector<int> dest;        // (K0) stored channel of each given one
  std::vector<double> coef;     // (K0) its grid's coefficient there
  std::vector<int> sign;        // (K0) (-1)^n_k against its leader
  std::vector<char> linked;     // (K0) member of a folded group
  std::vector<double> amp, off; // (K0) the pack's a_k, b_k
  std::vector<int> leader;      // (K1) given channel a stored one flies,
                                // -1 for the constant set
  int const_ch = -1;            // stored index of the constant set
};

struct Fields {
  FieldsK k{};
  DevBuf grid, metal, label, tab_t, tab_v, waves;
  Waves waves_host{};
  FoldPlan fold;
  int64_t bytes = 0;
};

// per-electrode bases for composition on the device
struct Bases {
  int32_t route = IG_ROUTE_RZ;
  int64_t n0 = 0, n1 = 0, n2 = 0;
  int32_t n_bases = 0;
  double base_volts = 1.0;
  DevBuf phi, metal, abc0, abc1, work, terms_i, terms_w;
  GradAxis g0{}, g1{};
};

// a captured loop of flight windows and the kernel arguments it replays
struct GraphEntry {
  std::vector<uint8_t> key;
  hipGraphExec_t exec;
};
constexpr size_t kMaxGraphs = 16;

struct Ensemble {
  int64_t cap = 0, n = 0;
  int32_t route = IG_ROUTE_RZ;
  // units of the state: -1 not flown yet; 0 t in us and v in mm/us (r-z,
  // planar, 3-D SDS); 1 t in s and v in m/s once flown (3-D RK4)
  int32_t clock = -1;
  int32_t rec_rows = 0, rec_cols = 0;
  int32_t n_tunes = 0;
  DevBuf tunes;         // Tune table (device)
  std::vector<ig_tune> tunes_in; // the table as given (host)
  DevBuf tunes_fold;    // the table re-laid for one folded pack (device)
  int64_t fold_for = 0; // the pack tunes_fold holds the table for (0: none)
  std::vector<Tune> fold_host; // staging of tunes_fold's last copy
  DevBuf counts;        // live_by_tune scratch (device)
  DevBuf sums;          // live/steps reduction (device)
  DevBuf live_idx;      // compacted live-ion list (int32, capacity)
  DevBuf live_cnt;      // its length (uint32)
  DevBuf live_key, live_sorted, live_hist; // its cell order (compact 2)
  int32_t *tune_idx = nullptr;
  std::vector<std::unique_ptr<DevBuf>> bufs;
  EnsK k{};
  template <class T>
  int alloc(T *&ptr, size_t count) {
    bufs.emplace_back(new DevBuf());
    int rc = dev_alloc(*bufs.back(), count * sizeof(T));
    ptr = (T *)bufs.back()->p;
    return rc;
  }
};

// ---- scratch of the no-allocation flight (ig_fly_noalloc) -------------
// One fixed region of device memory per GPU, part of the loaded code
// object: the live-ion list, its cell order, the counting-sort histogram
// and the SDS jump table of ig_fly_noalloc.  In host code `scratch` names a
// host-side stand-in, never device memory: each device's address comes from
// hipGetSymbolAddress, taken once per device (with that device current) when
// an engine is created on it, and only that address is passed to kernels
// and copies.  Calls of ig_fly_noalloc on one device share the region, so
// calls that overlap in time on one device race; each call rebuilds the
// list and histogram it reads (the counter and histogram are cleared by
// kernels on the call's stream) and rewrites the jump table when it
// differs from the last one copied.
constexpr int64_t kNoAllocMaxIons = IG_NOALLOC_MAX_IONS;
constexpr size_t kNaAlign = 256;
constexpr size_t na_round(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaHistWords = (1u << 18) + 256; // cell buckets + scan tile sums
constexpr size_t kNaSdsDoubles = 5 * 1002;
constexpr size_t kNaIdx = kNaAlign;                // [0, 4): the live count
constexpr size_t kNaKey = kNaIdx + na_round((size_t)kNoAllocMaxIons * 4);
constexpr size_t kNaSorted = kNaKey + na_round((size_t)kNoAllocMaxIons * 4);
constexpr size_t kNaHist = kNaSorted + na_round((size_t)kNoAllocMaxIons * 4);
constexpr size_t kNaSds = kNaHist + na_round(kNaHistWords * 4);
constexpr size_t kNaBytes = kNaSds + na_round(kNaSdsDoubles * 8);

// what one device's region holds, as the host knows it
struct NoAllocDevice {
  uint8_t *ptr = nullptr;       // hipGetSymbolAddress of scratch
  int status = IG_E_STATE;      // its return code, as an IG_E_* code
  std::string err;
  std::vector<double> sds_host; // the jump table last copied in
};
std::mutex g_na_mu;
std::map<int, NoAllocDevice> g_na; // by device ordinal; nodes never move

} // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {
// the device's region, looked up once per device with the device current
void noalloc_setup(int device) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  NoAllocDevice &d = g_na[device];
  if (d.ptr) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p) {
    d.ptr = (uint8_t *)p;
    d.status = 0;
  } else {
    d.status = IG_E_HIP;
    d.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
  }
}
} // namespace

struct ig_engine {
  int device = 0;
  DevBuf sds;                   // SDS jump table on the device
  std::vector<double> sds_host; // its host copy, to detect a change
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  int64_t next_id = 1;
  std::map<int64_t, std::unique_ptr<Fields>> fields;
  std::map<int64_t, std::unique_ptr<Ensemble>> ens;
  std::map<int64_t, std::unique_ptr<Bases>> bases;
  int block = 256;
  int cus = 256; // compute units of the device
  std::vector<GraphEntry> graphs; // captured window loops (ig_fly_graph)
  DevBuf bw, bj, bflag, btab, bout; // births: words, jumps, status, tables, rows
};

namespace {
// lanes per workgroup of a flight launch: the caller's, else the engine's
// (256) for an ensemble of at least eight 64-lane waves per compute unit,
// and 64 below that, so that a small ensemble's waves spread evenly over
// every compute unit instead of filling some with whole 256-lane
// workgroups (a 100,000-ion 3-D FP64 window: 49 ms against 57); results
// do not depend on it
int launch_block(const ig_engine *h, const ig_flight_params *p, int64_t n) {
  if (p->block > 0) return p->block;
  return (n + 63) / 64 < 8 * (int64_t)h->cus ? 64 : h->block;
}

void drop_graphs(ig_engine *h) {
  for (auto &g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
}
} // namespace

namespace {

int set_dev(ig_engine *h) {
  HIPX(hipSetDevice(h->device));
  return 0;
}

// copy host data into a fresh device buffer on the engine's stream
int upload(ig_engine *h, DevBuf &b, const void *src, size_t bytes) {
  int rc = dev_alloc(b, bytes);
  if (rc) return rc;
  if (bytes) HIPX(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, h->stream));
  return 0;
}

} // namespace

extern "C" {

const char *ig_last_error(void) { return g_err.c_str(); }

int ig_abi_version(void) { return IG_ABI_VERSION; }

int ig_device_count(int *out) {
  if (!out) return fail(IG_E_ARG, "ig_device_count: null out");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *out = n;
  return 0;
}

int ig_device_info(int device, char *name, int32_t name_len, int32_t *cu_count) {
  if (!name || name_len < 1 || !cu_count) return fail(IG_E_ARG, "ig_device_info: null argument");
  hipDeviceProp_t pr;
  HIPX(hipGetDeviceProperties(&pr, device));
  std::snprintf(name, (size_t)name_len, "%s %s", pr.name, pr.gcnArchName);
  *cu_count = pr.multiProcessorCount;
  return 0;
}

int ig_engine_create(int device, ig_engine **out) {
  if (!out) return fail(IG_E_ARG, "ig_engine_create: null out");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(IG_E_HIP, "ig_engine_create: no HIP device is visible");
  if (device < 0 || device >= n)
    return fail(IG_E_ARG, "ig_engine_create: device " + std::to_string(device) + " not in [0, " +
                              std::to_string(n) + ")");
  std::unique_ptr<ig_engine> h(new ig_engine());
  h->device = device;
  HIPX(hipSetDevice(device));
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) ==
            hipSuccess &&
        cus > 0)
      h->cus = cus;
  }
  noalloc_setup(device); // an error is kept and reported by ig_fly_noalloc
  HIPX(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPX(hipEventCreate(&h->ev0));
  HIPX(hipEventCreate(&h->ev1));
  *out = h.release();
  return 0;
}

int ig_engine_destroy(ig_engine *h) {
  if (!h) return 0;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  drop_graphs(h);
  h->fields.clear();
  h->ens.clear();
  h->bases.clear();
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int ig_sync(ig_engine *h) {
  if (!h) return fail(IG_E_ARG, "ig_sync: null engine");
  if (int rc = set_dev(h)) return rc;
  HIPX(hipStreamSynchronize(h->stream));
  return 0;
}

// a waveform w with w(x + pi) = -w(x)
static bool linkable(int kind, double duty) {
  return kind == IG_WAVE_SIN || kind == IG_WAVE_COS || (kind == IG_WAVE_SQUARE && duty == 0.5);
}

// dph = n pi for a whole n (to 1e-9 of a half period)
static bool half_periods(double dph, int &n) {
  if (!std::isfinite(dph)) return false;
  const double q = dph / 3.141592653589793;
  if (!(std::fabs(q) < 1.0e9)) return false;
  const double r = std::nearbyint(q);
  if (std::fabs(q - r) > 1.0e-9) return false;
  n = (int)(int64_t)r;
  return true;
}

static bool close_rel(double x, double y) {
  return std::fabs(x - y) <= 1.0e-9 * std::max(std::fabs(x), std::fabs(y));
}

// whether a flight kernel is instantiated for n_sets field sets on the
// route in the FP32-basis mode (the launchers' set-count switches)
static bool f32_kernel_for(int route, int n_sets) {
  if (route == IG_ROUTE_RZ) return n_sets >= 1 && n_sets <= 4;
  if (route == IG_ROUTE_PLANAR) return n_sets >= 1 && n_sets <= 5;
  return n_sets == 1 || n_sets == 2 || n_sets == 3 || n_sets == 7 || n_sets == 11;
}

// The fold of a pack's drive channels, or an inactive plan where folding is
// not asked for, not allowed (FP64 arithmetic, table waveforms), would not
// lower the set count or would give a set count no kernel is built for.
static FoldPlan make_fold_plan(const ig_fields_desc *d, bool affine) {
  FoldPlan P;
  const int K = d->n_chan;
  P.affine = affine;
  P.K0 = K;
  if (d->fold_channels != IG_FOLD_LINKED || d->precision != IG_PREC_F32 || K < 2) return P;
  for (int k = 0; k < K; ++k)
    if (d->ch_kind[k] == IG_WAVE_TAB_HOLD || d->ch_kind[k] == IG_WAVE_TAB_LINEAR) return P;
  P.amp.assign(K, 1.0);
  P.off.assign(K, 0.0);
  if (affine)
    for (int k = 0; k < K; ++k) {
      P.amp[k] = d->ch_amp ? d->ch_amp[k] : 1.0;
      P.off[k] = d->ch_off ? d->ch_off[k] : 0.0;
    }
  std::vector<int> lead(K);
  P.sign.assign(K, 1);
  for (int k = 0; k < K; ++k) {
    lead[k] = k;
    const bool ok = linkable(d->ch_kind[k], d->ch_duty[k]) && std::isfinite(d->ch_om[k]) &&
                    std::isfinite(P.amp[k]) && P.amp[k] != 0.0;
    if (!ok) continue;
    for (int L = 0; L < k; ++L) {
      int n = 0;
      if (lead[L] != L || !linkable(d->ch_kind[L], d->ch_duty[L]) || P.amp[L] == 0.0 ||
          !std::isfinite(P.amp[L]) || d->ch_kind[L] != d->ch_kind[k] ||
          d->ch_om[L] != d->ch_om[k] || !half_periods(d->ch_ph[k] - d->ch_ph[L], n))
        continue;
      lead[k] = L;
      P.sign[k] = (n & 1) ? -1 : 1;
      break;
    }
  }
  std::vector<int> members(K, 0);
  for (int k = 0; k < K; ++k) members[lead[k]] += 1;
  P.linked.assign(K, 0);
  bool any_off = false;
  for (int k = 0; k < K; ++k) {
    P.linked[k] = members[lead[k]] >= 2;
    any_off = any_off || (P.linked[k] && P.off[k] != 0.0);
  }
  int K1 = any_off ? 1 : 0;
  for (int k = 0; k < K; ++k) K1 += lead[k] == k;
  const int gate = d->route == IG_ROUTE_RZ && d->has_gate ? 1 : 0;
  if (K1 >= K || !f32_kernel_for(d->route, 1 + K1 + gate)) return P;
  P.dest.assign(K, -1);
  P.coef.assign(K, 1.0);
  for (int k = 0; k < K; ++k) {
    if (lead[k] == k) {
      P.dest[k] = (int)P.leader.size();
      P.leader.push_back(k);
    } else {
      P.dest[k] = P.dest[lead[k]];
    }
    if (P.linked[k]) P.coef[k] = P.sign[k] * (P.amp[k] / P.amp[lead[k]]);
  }
  if (any_off) {
    P.const_ch = (int)P.leader.size();
    P.leader.push_back(-1);
  }
  P.K1 = (int)P.leader.size();
  P.on = true;
  return P;
}

// A tune's drive re-laid onto a folded pack's stored channels.  The tune
// must keep every folded group linked as the pack has it: the same kind
// and frequency, phases the same parity of half periods apart, effective
// amplitudes (amplitude x set scale; the set scale alone on the unit-
// waveform routes) in the pack's ratios, and offsets one common multiple
// of the pack's.  Returns an empty string, or what breaks the fold.
static std::string fold_tune(const FoldPlan &P, const ig_tune &a, ig_tune &o) {
  std::memset(&o, 0, sizeof(o));
  o.dt = a.dt;
  o.t_max = a.t_max;
  o.rng_key = a.rng_key;
  for (int s = 0; s < IG_MAX_CHANNELS + 2; ++s) o.set_scale[s] = 1.0;
  for (int c = 0; c < IG_MAX_CHANNELS; ++c) {
    o.ch_kind[c] = IG_WAVE_SIN;
    o.ch_duty[c] = 0.5;
    o.ch_amp[c] = 1.0;
  }
  o.set_scale[0] = a.set_scale[0];
  o.set_scale[1 + P.K1] = a.set_scale[1 + P.K0]; // the r-z gate set
  auto alpha = [&](int k) { return P.affine ? a.ch_amp[k] * a.set_scale[1 + k] : a.set_scale[1 + k]; };
  auto beta = [&](int k) { return P.affine ? a.ch_off[k] * a.set_scale[1 + k] : 0.0; };
  // a channel whose weight is identically zero (no amplitude, or sin of a
  // constant zero phase) and no offset: a group of only such channels
  // (an idle tune row, say) stays zero whatever its phases
  auto zero_drive = [&](int k) {
    const bool still = a.ch_kind[k] == IG_WAVE_SIN && a.ch_om[k] == 0.0 && a.ch_ph[k] == 0.0;
    return (alpha(k) == 0.0 || still) && beta(k) == 0.0;
  };
  std::vector<char> zero_group(P.K1, 1);
  for (int k = 0; k < P.K0; ++k)
    if (P.linked[k] && !zero_drive(k)) zero_group[P.dest[k]] = 0;
  for (int j = 0; j < P.K1; ++j) {
    const int L = P.leader[j];
    if (L < 0) continue;
    if (P.linked[L] && zero_group[j]) { // sin(0) = 0: weight 0
      o.ch_amp[j] = 0.0;
      continue;
    }
    o.ch_kind[j] = a.ch_kind[L];
    o.ch_om[j] = a.ch_om[L];
    o.ch_ph[j] = a.ch_ph[L];
    o.ch_duty[j] = a.ch_duty[L];
    o.ch_amp[j] = a.ch_amp[L];
    o.ch_off[j] = P.linked[L] ? 0.0 : a.ch_off[L];
    o.set_scale[1 + j] = a.set_scale[1 + L];
  }
  double mu = 0.0;
  bool mu_set = false;
  for (int k = 0; k < P.K0; ++k) {
    if (!P.linked[k]) continue;
    const int L = P.leader[P.dest[k]];
    const std::string ch = "channel " + std::to_string(k) + " ";
    if (zero_group[P.dest[k]]) {
      if (P.off[k] != 0.0 && P.const_ch >= 0) { // its offset share is 0
        if (!mu_set) {
          mu = 0.0;
          mu_set = true;
        } else if (mu != 0.0) {
          return ch + "scales its offset unlike the other linked channels";
        }
      }
      continue;
    }
    if (!linkable(a.ch_kind[k], a.ch_duty[k]) || a.ch_kind[k] != a.ch_kind[L] ||
        a.ch_duty[k] != a.ch_duty[L])
      return ch + "changes its waveform kind away from channel " + std::to_string(L) + "'s";
    if (a.ch_om[k] != a.ch_om[L])
      return ch + "changes its frequency away from channel " + std::to_string(L) + "'s";
    int n = 0;
    if (!half_periods(a.ch_ph[k] - a.ch_ph[L], n) || ((n & 1) ? -1 : 1) != P.sign[k])
      return ch + "changes its phase against channel " + std::to_string(L) + "'s";
    if (!close_rel(alpha(k) * P.amp[L], alpha(L) * P.amp[k]))
      return ch + "changes its amplitude against channel " + std::to_string(L) + "'s";
    const double b = beta(k);
    if (P.off[k] == 0.0) {
      if (b != 0.0) return ch + "gains an offset the pack does not have";
    } else {
      const double m = b / P.off[k];
      if (!mu_set) {
        mu = m;
        mu_set = true;
      } else if (!close_rel(m, mu)) {
        return ch + "scales its offset unlike the other linked channels";
      }
    }
  }
  if (P.const_ch >= 0) {
    const int j = P.const_ch;
    o.ch_kind[j] = IG_WAVE_SIN; // sin(0) = 0: the weight is the offset
    o.ch_om[j] = 0.0;
    o.ch_ph[j] = 0.0;
    o.ch_duty[j] = 0.5;
    o.ch_amp[j] = 0.0;
    o.ch_off[j] = mu;
    o.set_scale[1 + j] = 1.0;
  }
  return std::string();
}

int ig_fields_upload(ig_engine *h, const ig_fields_desc *d, ig_fields_id *out) {
  if (!h || !d || !out) return fail(IG_E_ARG, "ig_fields_upload: null argument");
  if (d->route != IG_ROUTE_RZ && d->route != IG_ROUTE_XYZ && d->route != IG_ROUTE_PLANAR)
    return fail(IG_E_ARG, "ig_fields_upload: unknown route " + std::to_string(d->route));
  if (d->precision < IG_PREC_F64 || d->precision > IG_PREC_F32)
    return fail(IG_E_ARG, "ig_fields_upload: unknown precision");
  const bool rz = d->route == IG_ROUTE_RZ;
  const bool two_d = d->route != IG_ROUTE_XYZ; // r-z and planar packs
  if (d->n0 < 2 || d->n1 < 2 || (two_d ? d->n2 != 1 : d->n2 < 2))
    return fail(IG_E_ARG, "ig_fields_upload: grid shape must be at least 2 nodes per axis "
                          "(n2 == 1 for the r-z and planar routes)");
  if (d->n0 >= INT32_MAX || d->n1 >= INT32_MAX || d->n2 >= INT32_MAX)
    return fail(IG_E_ARG, "ig_fields_upload: a grid axis of 2^31 or more nodes (the kernels "
                          "index nodes in 32 bits)");
  if (d->n_chan < 0 || d->n_chan > IG_MAX_CHANNELS)
    return fail(IG_E_ARG, "ig_fields_upload: n_chan out of [0, " +
                              std::to_string(IG_MAX_CHANNELS) + "]");
  if (!rz && d->has_gate) return fail(IG_E_ARG, "ig_fields_upload: the gate set is r-z only");
  if (!d->grids || !d->metal) return fail(IG_E_ARG, "ig_fields_upload: grids and metal required");
  if (!(d->h_mm > 0.0)) return fail(IG_E_ARG, "ig_fields_upload: h_mm must be > 0");
  if (d->n_chan > 0 && (!d->ch_kind || !d->ch_om || !d->ch_ph || !d->ch_duty))
    return fail(IG_E_ARG, "ig_fields_upload: waveform arrays required for drive channels");
  if (int rc = set_dev(h)) return rc;

  auto F = std::make_unique<Fields>();
  FieldsK &k = F->k;
  k.route = d->route;
  k.prec = d->precision;
  k.n0 = d->n0;
  k.n1 = d->n1;
  k.n2 = d->n2;
  {
    const int64_t nn[3] = {k.n0, k.n1, k.n2};
    for (int a = 0; a < 3; ++a) {
      k.dn[a] = (double)nn[a];
      k.dnm1[a] = (double)(nn[a] - 1);
      k.dnm2[a] = (double)(nn[a] - 2);
    }
  }
  k.h = d->h_mm;
  k.inv_h = 1.0 / d->h_mm;
  k.u0 = d->u0;
  k.has_gate = rz && d->has_gate ? 1 : 0;
  k.tau_gate = k.has_gate ? d->tau_gate : -1.0;
  const int n_sets_in = 1 + d->n_chan + k.has_gate;
  for (int s = 0; s < n_sets_in; ++s) {
    double sc = d->chan_scale ? d->chan_scale[s] : 1.0;
    if (!(sc > 0.0) || !std::isfinite(sc))
      return fail(IG_E_ARG, "ig_fields_upload: chan_scale must be finite and > 0");
  }
  for (int c = 0; c < d->n_chan; ++c) {
    int kind = d->ch_kind[c];
    if (kind < IG_WAVE_SIN || kind > IG_WAVE_TAB_LINEAR)
      return fail(IG_E_ARG, "ig_fields_upload: unknown waveform kind " + std::to_string(kind));
  }
  if (d->fold_channels != IG_FOLD_NONE && d->fold_channels != IG_FOLD_LINKED)
    return fail(IG_E_ARG, "ig_fields_upload: fold_channels must be IG_FOLD_NONE or IG_FOLD_LINKED");
  // the stored sets: as given, or with linked drive channels folded
  F->fold = make_fold_plan(d, !two_d);
  const FoldPlan &P = F->fold;
  k.K = P.on ? P.K1 : d->n_chan;
  k.n_sets = 1 + k.K + k.has_gate;
  k.n_comp = two_d ? 2 : 3;
  const int nvals = k.n_sets * k.n_comp;
  k.stride = record_stride(k.prec, nvals);
  // the FP32-basis 2-D kernels address records and metal nodes with 32-bit
  // offsets
  if (two_d && k.prec == IG_PREC_F32 &&
      ((double)d->n0 * (double)d->n1 >= 2147483648.0 ||
       (double)d->n0 * (double)d->n1 * (double)k.stride * 4.0 >= 4294967296.0))
    return fail(IG_E_ARG, "ig_fields_upload: an FP32-basis 2-D pack of 2^31 nodes or 4 GiB of "
                          "records or more (the kernels' 32-bit offsets)");
  // given set of each stored set (-1: a folded drive set), and the stored
  // storage scale: the leader's for a folded group, 1 for the constant set
  std::vector<int> given((size_t)k.n_sets, -1);
  given[0] = 0;
  for (int j = 0; j < k.K; ++j) given[1 + j] = P.on ? -1 : 1 + j;
  if (k.has_gate) given[1 + k.K] = 1 + d->n_chan;
  for (int s = 0; s < k.n_sets; ++s) {
    int src = given[s];
    if (src < 0) src = P.leader[s - 1] >= 0 ? 1 + P.leader[s - 1] : -1;
    const double sc = src >= 0 && d->chan_scale ? d->chan_scale[src] : 1.0;
    k.scale[s] = k.prec == IG_PREC_F64 ? 1.0 : sc;
  }
  int64_t tab_total = 0;
  Waves &wv = F->waves_host;
  std::memset(&wv, 0, sizeof(wv));
  for (int c = 0; c < k.K; ++c) {
    const int L = P.on ? P.leader[c] : c;
    if (L < 0) { // the constant set: sin(0 t + 0) = 0, weight 0 * 0 + 1
      wv.kind[c] = IG_WAVE_SIN;
      wv.duty[c] = 0.5;
      wv.amp[c] = 0.0;
      wv.off[c] = 1.0;
      continue;
    }
    wv.kind[c] = d->ch_kind[L];
    wv.om[c] = d->ch_om[L];
    wv.ph[c] = d->ch_ph[L];
    wv.duty[c] = d->ch_duty[L];
    wv.amp[c] = d->ch_amp ? d->ch_amp[L] : 1.0;
    wv.off[c] = d->ch_off ? d->ch_off[L] : 0.0;
    if (P.on && P.linked[L]) wv.off[c] = 0.0; // carried by the constant set
  }
  // (a folded pack has no table waveforms)
  for (int c = 0; c <= k.K; ++c) wv.tab_off[c] = d->tab_off && !P.on ? d->tab_off[c] : 0;
  if (d->tab_off && !P.on) {
    tab_total = d->n_tab;
    for (int c = 0; c < k.K; ++c)
      if (wv.tab_off[c] < 0 || wv.tab_off[c + 1] < wv.tab_off[c] || wv.tab_off[c + 1] > tab_total)
        return fail(IG_E_ARG, "ig_fields_upload: tab_off slices out of range");
  }
  for (int c = 0; c < k.K; ++c)
    if ((wv.kind[c] == IG_WAVE_TAB_HOLD || wv.kind[c] == IG_WAVE_TAB_LINEAR) &&
        wv.tab_off[c + 1] > wv.tab_off[c] && (!d->tab_t || !d->tab_v))
      return fail(IG_E_ARG, "ig_fields_upload: table channel without tab_t/tab_v");

  // interleave (set, comp) per node into padded records
  const int64_t nodes = d->n0 * d->n1 * d->n2;
  const int64_t S = k.stride;
  const bool f32 = k.prec != IG_PREC_F64;
  std::vector<double> h64;
  std::vector<float> h32;
  if (f32)
    h32.assign((size_t)(nodes * S), 0.0f);
  else
    h64.assign((size_t)(nodes * S), 0.0);
  std::vector<double> comb;
  for (int s = 0; s < k.n_sets; ++s) {
    const double inv = 1.0 / (k.prec == IG_PREC_F64 ? 1.0 : k.scale[s]);
    for (int c = 0; c < k.n_comp; ++c) {
      const double *src;
      if (given[s] >= 0) {
        src = d->grids + ((int64_t)given[s] * k.n_comp + c) * nodes;
      } else {
        // a folded set, summed in FP64 in the given channels' order
        const int j = s - 1;
        comb.assign((size_t)nodes, 0.0);
        for (int g = 0; g < P.K0; ++g) {
          double w;
          if (P.leader[j] >= 0) {
            if (P.dest[g] != j) continue;
            w = P.coef[g];
          } else {
            if (!P.linked[g] || P.off[g] == 0.0) continue;
            w = P.off[g];
          }
          const double *e = d->grids + ((int64_t)(1 + g) * k.n_comp + c) * nodes;
          for (int64_t n = 0; n < nodes; ++n) comb[n] += w * e[n];
        }
        src = comb.data();
      }
      const int q = s * k.n_comp + c;
      if (f32) {
        for (int64_t n = 0; n < nodes; ++n) h32[n * S + q] = (float)(src[n] * inv);
      } else {
        for (int64_t n = 0; n < nodes; ++n) h64[n * S + q] = src[n];
      }
    }
  }
  int rc = f32 ? upload(h, F->grid, h32.data(), h32.size() * sizeof(float))
               : upload(h, F->grid, h64.data(), h64.size() * sizeof(double));
  if (rc) return rc;
  std::vector<uint8_t> m((size_t)nodes);
  for (int64_t n = 0; n < nodes; ++n) m[n] = d->metal[n] ? 1 : 0;
  if ((rc = upload(h, F->metal, m.data(), m.size()))) return rc;
  if (d->label && (rc = upload(h, F->label, d->label, (size_t)nodes))) return rc;
  if (tab_total > 0) {
    if ((rc = upload(h, F->tab_t, d->tab_t, tab_total * sizeof(double)))) return rc;
    if ((rc = upload(h, F->tab_v, d->tab_v, tab_total * sizeof(double)))) return rc;
  }
  if ((rc = upload(h, F->waves, &F->waves_host, sizeof(Waves)))) return rc;
  HIPX(hipStreamSynchronize(h->stream));
  k.grid = F->grid.p;
  k.metal = (const uint8_t *)F->metal.p;
  k.label = (const uint8_t *)F->label.p;
  k.tab_t = (const double *)F->tab_t.p;
  k.tab_v = (const double *)F->tab_v.p;
  k.waves = (const Waves *)F->waves.p;
  F->bytes = (int64_t)(F->grid.bytes + F->metal.bytes + F->label.bytes + F->tab_t.bytes +
                       F->tab_v.bytes);
  int64_t id = h->next_id++;
  h->fields[id] = std::move(F);
  *out = id;
  return 0;
}

int ig_fields_release(ig_engine *h, ig_fields_id id) {
  if (!h) return fail(IG_E_ARG, "ig_fields_release: null engine");
  auto it = h->fields.find(id);
  if (it == h->fields.end()) return fail(IG_E_STATE, "ig_fields_release: unknown field id");
  if (int rc = set_dev(h)) return rc;
  HIPX(hipStreamSynchronize(h->stream));
  drop_graphs(h);
  h->fields.erase(it);
  return 0;
}

int ig_fields_bytes(ig_engine *h, ig_fields_id id, int64_t *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_fields_bytes: null argument");
  auto it = h->fields.find(id);
  if (it == h->fields.end()) return fail(IG_E_STATE, "ig_fields_bytes: unknown field id");
  *out = it->second->bytes;
  return 0;
}

int ig_ensemble_create(ig_engine *h, int64_t capacity, int32_t rec_rows, int32_t rec_cols,
                       ig_ensemble_id *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_ensemble_create: null argument");
  if (capacity < 0 || rec_rows < 0 || rec_cols < 0 || (rec_rows > 0 && rec_cols < 7))
    return fail(IG_E_ARG, "ig_ensemble_create: bad capacity or record shape");
  if (int rc = set_dev(h)) return rc;
  auto E = std::make_unique<Ensemble>();
  E->cap = capacity;
  E->rec_rows = rec_rows;
  E->rec_cols = rec_cols;
  const size_t n = (size_t)(capacity > 0 ? capacity : 1);
  EnsK &k = E->k;
  int rc = 0;
  double **f64[] = {&k.x,     &k.y,        &k.z,       &k.vx,    &k.vy,       &k.vz,
                    &k.t,     &k.tob,      &k.m_amu,   &k.acc,   &k.sigma,    &k.sds_damp,
                    &k.sds_mfp, &k.sds_V,  &k.sds_logmr, &k.gauss, &k.path,   &k.eat,
                    &k.ket};
  for (double **a : f64)
    if ((rc = E->alloc(*a, n))) return rc;
  if ((rc = E->alloc(k.e_last, 3 * n))) return rc;
  if ((rc = E->alloc(k.seed, n))) return rc;
  if ((rc = E->alloc(k.rng_n, n))) return rc;
  if ((rc = E->alloc(k.ncol, n))) return rc;
  if ((rc = E->alloc(k.steps, n))) return rc;
  if ((rc = E->alloc(k.hz, n))) return rc;
  int32_t **i32[] = {&k.has_gauss, &k.fate, &k.face, &k.label, &k.nrec, &k.tp_pass,
                     &E->tune_idx};
  for (int32_t **a : i32)
    if ((rc = E->alloc(*a, n))) return rc;
  // slots hold no ion until loaded: a fate the kernels skip
  {
    std::vector<int32_t> empty(n, IG_FATE_EMPTY);
    HIPX(hipMemcpy(k.fate, empty.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPX(hipMemset(E->tune_idx, 0, n * sizeof(int32_t)));
  }
  k.tune = E->tune_idx;
  k.tunes = nullptr;
  k.rec = nullptr;
  if (rec_rows > 0 && (rc = E->alloc(k.rec, n * (size_t)rec_rows * rec_cols))) return rc;
  k.idx = nullptr;
  k.n_live = nullptr;
  k.n = 0;
  int64_t id = h->next_id++;
  h->ens[id] = std::move(E);
  *out = id;
  return 0;
}

static Ensemble *find_ens(ig_engine *h, ig_ensemble_id id) {
  auto it = h->ens.find(id);
  return it == h->ens.end() ? nullptr : it->second.get();
}

int ig_ensemble_release(ig_engine *h, ig_ensemble_id id) {
  if (!h) return fail(IG_E_ARG, "ig_ensemble_release: null engine");
  auto it = h->ens.find(id);
  if (it == h->ens.end()) return fail(IG_E_STATE, "ig_ensemble_release: unknown ensemble id");
  if (int rc = set_dev(h)) return rc;
  HIPX(hipStreamSynchronize(h->stream));
  drop_graphs(h);
  h->ens.erase(it);
  return 0;
}

static int load_range(ig_engine *h, Ensemble *E, int64_t off, const ig_ions *io, int32_t tune,
                      const char *who) {
  const std::string w(who);
  if (io->n < 0 || off < 0 || off + io->n > E->cap)
    return fail(IG_E_ARG, w + ": ions [" + std::to_string(off) + ", " +
                              std::to_string(off + io->n) + ") exceed the ensemble capacity " +
                              std::to_string(E->cap));
  const bool no_births =
      !io->x && !io->y && !io->z && !io->vx && !io->vy && !io->vz && !io->tob;
  if (io->n > 0 && ((!no_births && (!io->x || !io->y || !io->z || !io->vx || !io->vy ||
                                    !io->vz || !io->tob)) ||
                    !io->m_amu || !io->acc || !io->seed))
    return fail(IG_E_ARG, w + ": births (all seven or none), m_amu, acc and seed are required");
  if (tune < 0 || tune >= (E->n_tunes > 0 ? E->n_tunes : 1))
    return fail(IG_E_ARG, w + ": tune " + std::to_string(tune) + " is not in the tune table");
  if (int rc = set_dev(h)) return rc;
  const int64_t n = io->n;
  if (n == 0) return 0;
  EnsK &k = E->k;
  const size_t b8 = (size_t)n * 8, b4 = (size_t)n * 4;
  hipStream_t st = h->stream;
  struct {
    void *dst;
    const void *src;
  } cp[] = {{k.x + off, io->x},         {k.y + off, io->y},     {k.z + off, io->z},
            {k.vx + off, io->vx},       {k.vy + off, io->vy},   {k.vz + off, io->vz},
            {k.tob + off, io->tob},     {k.m_amu + off, io->m_amu}, {k.acc + off, io->acc},
            {k.seed + off, io->seed}};
  for (auto &c : cp) {
    if (c.src)
      HIPX(hipMemcpyAsync(c.dst, c.src, b8, hipMemcpyHostToDevice, st));
    else
      HIPX(hipMemsetAsync(c.dst, 0, b8, st));
  }
  if (io->sigma_m2)
    HIPX(hipMemcpyAsync(k.sigma + off, io->sigma_m2, b8, hipMemcpyHostToDevice, st));
  else
    HIPX(hipMemsetAsync(k.sigma + off, 0, b8, st));
  double *sds_cols[4] = {k.sds_damp + off, k.sds_mfp + off, k.sds_V + off, k.sds_logmr + off};
  std::vector<double> col;
  for (int c = 0; c < 4; ++c) {
    if (io->sds) {
      // (n, 4) rows -> four columns
      col.resize((size_t)n);
      for (int64_t i = 0; i < n; ++i) col[i] = io->sds[4 * i + c];
      HIPX(hipMemcpyAsync(sds_cols[c], col.data(), b8, hipMemcpyHostToDevice, st));
      HIPX(hipStreamSynchronize(st));
    } else {
      HIPX(hipMemsetAsync(sds_cols[c], 0, b8, st));
    }
  }
  void *zero8[] = {k.t + off,     k.rng_n + off, k.gauss + off, k.ncol + off,
                   k.steps + off, k.path + off,  k.eat + off,   k.ket + off};
  for (void *z : zero8) HIPX(hipMemsetAsync(z, 0, b8, st));
  HIPX(hipMemsetAsync(k.e_last + 3 * off, 0, 3 * b8, st));
  HIPX(hipMemsetAsync(k.has_gauss + off, 0, b4, st));
  HIPX(hipMemsetAsync(k.hz + off, 0xff, b4, st)); // NaN: no hazard drawn yet
  HIPX(hipMemsetAsync(k.face + off, 0xff, b4, st)); // -1
  HIPX(hipMemsetAsync(k.label + off, 0, b4, st));
  HIPX(hipMemsetAsync(k.nrec + off, 0, b4, st));
  HIPX(hipMemsetAsync(k.tp_pass + off, 0, b4, st));
  std::vector<int32_t> fz((size_t)n, IG_FATE_FLYING), tz((size_t)n, tune);
  if (io->fate_in)
    for (int64_t i = 0; i < n; ++i) fz[i] = io->fate_in[i];
  HIPX(hipMemcpyAsync(k.fate + off, fz.data(), b4, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(E->tune_idx + off, tz.data(), b4, hipMemcpyHostToDevice, st));
  HIPX(hipStreamSynchronize(st));
  return 0;
}

int ig_ensemble_load(ig_engine *h, ig_ensemble_id id, const ig_ions *io) {
  if (!h || !io) return fail(IG_E_ARG, "ig_ensemble_load: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_load: unknown ensemble id");
  int rc = load_range(h, E, 0, io, 0, "ig_ensemble_load");
  if (rc) return rc;
  E->clock = -1; // every ion replaced: not flown yet
  E->n = io->n;
  E->k.n = io->n;
  return 0;
}

int ig_ensemble_load_at(ig_engine *h, ig_ensemble_id id, int64_t offset, const ig_ions *io,
                        int32_t tune) {
  if (!h || !io) return fail(IG_E_ARG, "ig_ensemble_load_at: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_load_at: unknown ensemble id");
  int rc = load_range(h, E, offset, io, tune, "ig_ensemble_load_at");
  if (rc) return rc;
  if (offset + io->n > E->n) E->n = E->k.n = offset + io->n;
  return 0;
}

static int read_range(ig_engine *h, Ensemble *E, int64_t off, int64_t n, ig_results *o) {
  if (off < 0 || n < 0 || off + n > E->n)
    return fail(IG_E_ARG, "ig_ensemble_read: range [" + std::to_string(off) + ", " +
                              std::to_string(off + n) + ") is outside the loaded ions");
  if (int rc = set_dev(h)) return rc;
  const EnsK &k = E->k;
  if (n == 0) return 0;
  const size_t b8 = (size_t)n * 8, b4 = (size_t)n * 4;
  hipStream_t st = h->stream;
  HIPX(hipStreamSynchronize(st));
  struct {
    void *dst;
    const void *src;
    size_t bytes;
  } cp[] = {{o->x, k.x + off, b8},         {o->y, k.y + off, b8},
            {o->z, k.z + off, b8},         {o->vx, k.vx + off, b8},
            {o->vy, k.vy + off, b8},       {o->vz, k.vz + off, b8},
            {o->fate, k.fate + off, b4},   {o->face, k.face + off, b4},
            {o->label, k.label + off, b4}, {o->ncol, k.ncol + off, b8},
            {o->steps, k.steps + off, b8}, {o->nrec, k.nrec + off, b4}};
  for (auto &c : cp)
    if (c.dst) HIPX(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, st));
  const bool si = E->route == IG_ROUTE_XYZ && E->clock == 1;
  if (si && (o->vx || o->vy || o->vz)) {
    // the 3-D kernel carries m/s from its first entry on (nrec >= 1);
    // report mm/us as the reference does (tracer3d.py:779: v * 1e-3)
    std::vector<int32_t> nr((size_t)n);
    HIPX(hipMemcpyAsync(nr.data(), k.nrec + off, b4, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    double *vv[3] = {o->vx, o->vy, o->vz};
    for (double *v : vv)
      if (v)
        for (int64_t i = 0; i < n; ++i)
          if (nr[i] > 0) v[i] = v[i] * 1e-3;
  }
  if (o->t_us) {
    // lab time of the end state: tob + t (r-z, t in us) or tob + t*1e6
    // (3-D, t in s), evaluated on the host in the reference's order
    std::vector<double> t((size_t)n), tob((size_t)n);
    HIPX(hipMemcpyAsync(t.data(), k.t + off, b8, hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(tob.data(), k.tob + off, b8, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    if (si)
      for (int64_t i = 0; i < n; ++i) o->t_us[i] = tob[i] + t[i] * 1e6;
    else
      for (int64_t i = 0; i < n; ++i) o->t_us[i] = tob[i] + t[i];
  }
  if (o->records && E->rec_rows > 0 && k.rec)
    HIPX(hipMemcpyAsync(o->records, k.rec + (size_t)off * E->rec_rows * E->rec_cols,
                        (size_t)n * E->rec_rows * E->rec_cols * 8, hipMemcpyDeviceToHost, st));
  HIPX(hipStreamSynchronize(st));
  return 0;
}

int ig_ensemble_read(ig_engine *h, ig_ensemble_id id, ig_results *o) {
  if (!h || !o) return fail(IG_E_ARG, "ig_ensemble_read: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_read: unknown ensemble id");
  return read_range(h, E, 0, E->n, o);
}

int ig_ensemble_read_at(ig_engine *h, ig_ensemble_id id, int64_t offset, int64_t n,
                        ig_results *o) {
  if (!h || !o) return fail(IG_E_ARG, "ig_ensemble_read_at: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_read_at: unknown ensemble id");
  return read_range(h, E, offset, n, o);
}

// an ig_tune as the kernels read it
static int tune_from(const ig_tune &a, Tune &b, const std::string &who) {
  std::memset(&b, 0, sizeof(b));
  if (!(a.dt > 0.0) || !std::isfinite(a.dt)) return fail(IG_E_ARG, who + " dt <= 0");
  b.dt = a.dt;
  b.t_max = a.t_max;
  b.rng_key = a.rng_key;
  for (int s2 = 0; s2 < IG_MAX_CHANNELS + 2; ++s2) b.scale[s2] = a.set_scale[s2];
  for (int c = 0; c < IG_MAX_CHANNELS; ++c) {
    if (a.ch_kind[c] < IG_WAVE_SIN || a.ch_kind[c] > IG_WAVE_SQUARE)
      return fail(IG_E_ARG, who + " channel " + std::to_string(c) +
                                ": tunes carry sin, cos or square drives (tables stay with the "
                                "field pack)");
    b.w.kind[c] = a.ch_kind[c];
    b.w.om[c] = a.ch_om[c];
    b.w.ph[c] = a.ch_ph[c];
    b.w.duty[c] = a.ch_duty[c];
    b.w.amp[c] = a.ch_amp[c];
    b.w.off[c] = a.ch_off[c];
    WaveF &w = b.wf[c + 1];
    w.om = a.ch_om[c];
    w.ph = a.ch_ph[c];
    w.a = a.ch_amp[c] * a.set_scale[c + 1];
    w.b = a.ch_off[c] * a.set_scale[c + 1];
    const uint32_t mode =
        a.ch_kind[c] == IG_WAVE_SQUARE && a.ch_duty[c] != 0.5 ? kWaveSlow : a.ch_kind[c];
    b.wmode[(c + 1) >> 3] |= mode << (4 * ((c + 1) & 7));
  }
  b.wmode[0] |= (uint32_t)kWaveConst;
  b.wf[0].a = a.set_scale[0];
  return 0;
}

int ig_ensemble_set_tunes(ig_engine *h, ig_ensemble_id id, int32_t n, const ig_tune *tunes) {
  if (!h || (n > 0 && !tunes)) return fail(IG_E_ARG, "ig_ensemble_set_tunes: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_set_tunes: unknown ensemble id");
  if (n < 0 || n > (1 << 20)) return fail(IG_E_ARG, "ig_ensemble_set_tunes: bad tune count");
  if (int rc = set_dev(h)) return rc;
  HIPX(hipStreamSynchronize(h->stream));
  if (n == 0) {
    if (E->tunes.p) (void)hipFree(E->tunes.p);
    E->tunes.p = nullptr;
    E->tunes.bytes = 0;
    E->tunes_fold.reset();
    E->tunes_in.clear();
    E->fold_for = 0;
    E->n_tunes = 0;
    E->k.tunes = nullptr;
    return 0;
  }
  std::vector<Tune> host((size_t)n);
  for (int32_t t = 0; t < n; ++t)
    if (int rc = tune_from(tunes[t], host[t], "ig_ensemble_set_tunes: tune " + std::to_string(t)))
      return rc;
  // loaded ions must keep a valid tune index
  if (E->n > 0) {
    std::vector<int32_t> idx((size_t)E->n);
    HIPX(hipMemcpy(idx.data(), E->tune_idx, E->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int32_t v : idx)
      if (v >= n) return fail(IG_E_ARG, "ig_ensemble_set_tunes: a loaded ion uses tune " +
                                            std::to_string(v) + " beyond the new table");
  }
  DevBuf nb, nf;
  if (int rc = dev_alloc(nb, host.size() * sizeof(Tune))) return rc;
  // room for the table re-laid onto a folded pack's channels, filled when
  // the ensemble flies on one (so that no flight call allocates for it)
  if (int rc = dev_alloc(nf, host.size() * sizeof(Tune))) return rc;
  HIPX(hipMemcpy(nb.p, host.data(), host.size() * sizeof(Tune), hipMemcpyHostToDevice));
  std::swap(E->tunes.p, nb.p);
  std::swap(E->tunes.bytes, nb.bytes);
  std::swap(E->tunes_fold.p, nf.p);
  std::swap(E->tunes_fold.bytes, nf.bytes);
  E->tunes_in.assign(tunes, tunes + n);
  E->fold_for = 0;
  E->n_tunes = n;
  E->k.tunes = (const Tune *)E->tunes.p;
  return 0;
}

// The tune table the next flight of E on pack F reads: the table as given,
// or, on a pack whose linked channels were folded, the table re-laid onto
// the stored channels (built on the host and copied into the ensemble's
// second table when the pack differs from the last one it was built for).
static int bind_tunes(ig_engine *h, Ensemble *E, ig_fields_id fid, const Fields &F,
                      const std::string &who) {
  if (E->n_tunes == 0) return 0;
  if (!F.fold.on) {
    E->k.tunes = (const Tune *)E->tunes.p;
    return 0;
  }
  if (E->fold_for != fid) {
    std::vector<Tune> host((size_t)E->n_tunes);
    for (int32_t t = 0; t < E->n_tunes; ++t) {
      ig_tune o;
      const std::string why = fold_tune(F.fold, E->tunes_in[t], o);
      if (!why.empty())
        return fail(IG_E_ARG, who + ": tune " + std::to_string(t) + ": " + why +
                                  ", which the FP32-basis pack folded at upload "
                                  "(IG_FOLD_LINKED) cannot fly; upload it with IG_FOLD_NONE");
      if (int rc = tune_from(o, host[t], who + ": tune " + std::to_string(t))) return rc;
    }
    // an earlier copy from the staging vector may still be in flight
    HIPX(hipStreamSynchronize(h->stream));
    E->fold_host.swap(host);
    HIPX(hipMemcpyAsync(E->tunes_fold.p, E->fold_host.data(), E->fold_host.size() * sizeof(Tune),
                        hipMemcpyHostToDevice, h->stream));
    E->fold_for = fid;
  }
  E->k.tunes = (const Tune *)E->tunes_fold.p;
  return 0;
}

// live ions per tune slot: the lanes of a wave that share a tune add
// once, led by their first lane (a served slab keeps each request's ions
// contiguous, so a wave holds one or two tunes and the counters see one
// atomic per wave instead of one per ion)
__global__ void count_live_kernel(const int32_t *__restrict__ fate,
                                  const int32_t *__restrict__ tune, int64_t n, int32_t nt,
                                  unsigned long long *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t t = -1;
  if (i < n && fate[i] == IG_FATE_FLYING) {
    const int32_t v = tune[i];
    if (v >= 0 && v < nt) t = v;
  }
  uint64_t todo = __ballot(t >= 0);
  while (todo) {
    const int lead = __ffsll((unsigned long long)todo) - 1;
    const int32_t tl = __shfl(t, lead);
    const uint64_t same = __ballot(t == tl);
    if ((int)__lane_id() == lead) atomicAdd(&cnt[tl], (unsigned long long)__popcll(same));
    todo &= ~same;
  }
}

__global__ void zero_u64_kernel(unsigned long long *p, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

int ig_ensemble_live_by_tune(ig_engine *h, ig_ensemble_id id, int32_t n, int64_t *out) {
  if (!h || (n > 0 && !out)) return fail(IG_E_ARG, "ig_ensemble_live_by_tune: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_live_by_tune: unknown ensemble id");
  if (n <= 0) return 0;
  if (int rc = set_dev(h)) return rc;
  if (E->counts.bytes < (size_t)n * 8) {
    DevBuf nb;
    if (int rc = dev_alloc(nb, (size_t)n * 8)) return rc;
    HIPX(hipStreamSynchronize(h->stream));
    std::swap(E->counts.p, nb.p);
    std::swap(E->counts.bytes, nb.bytes);
  }
  auto *cnt = (unsigned long long *)E->counts.p;
  hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, h->stream, cnt, n);
  if (E->n > 0)
    hipLaunchKernelGGL(count_live_kernel, dim3((unsigned)((E->n + 255) / 256)), dim3(256), 0,
                       h->stream, E->k.fate, E->tune_idx, E->n, n, cnt);
  HIPX(hipGetLastError());
  HIPX(hipMemcpyAsync(out, cnt, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream));
  HIPX(hipStreamSynchronize(h->stream));
  return 0;
}

// live ions and the step total of an ensemble, reduced on the device:
// out[0] += live count, out[1] += sum of steps (grid-stride, one atomic
// pair per block)
__global__ void live_steps_kernel(const int32_t *__restrict__ fate,
                                  const int64_t *__restrict__ steps, int64_t n,
                                  unsigned long long *__restrict__ out) {
  __shared__ unsigned long long sl[256], ss[256];
  unsigned long long l = 0, st = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    l += fate[i] == IG_FATE_FLYING;
    st += (unsigned long long)steps[i];
  }
  sl[threadIdx.x] = l;
  ss[threadIdx.x] = st;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sl[threadIdx.x] += sl[threadIdx.x + w];
      ss[threadIdx.x] += ss[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], sl[0]);
    atomicAdd(&out[1], ss[0]);
  }
}

static int live_and_steps(ig_engine *h, Ensemble *E, int64_t *live, int64_t *steps) {
  if (int rc = set_dev(h)) return rc;
  if (E->sums.bytes < 16) {
    if (int rc = dev_alloc(E->sums, 16)) return rc;
  }
  auto *o = (unsigned long long *)E->sums.p;
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, h->stream, o, 2);
  if (E->n > 0) {
    const int64_t blocks = std::min<int64_t>((E->n + 255) / 256, 1024);
    hipLaunchKernelGGL(live_steps_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                       E->k.fate, E->k.steps, E->n, o);
  }
  HIPX(hipGetLastError());
  unsigned long long r[2] = {0, 0};
  HIPX(hipMemcpyAsync(r, o, 16, hipMemcpyDeviceToHost, h->stream));
  HIPX(hipStreamSynchronize(h->stream));
  if (live) *live = (int64_t)r[0];
  if (steps) *steps = (int64_t)r[1];
  return 0;
}

int ig_ensemble_live(ig_engine *h, ig_ensemble_id id, int64_t *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_ensemble_live: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_live: unknown ensemble id");
  return live_and_steps(h, E, out, nullptr);
}

int ig_ensemble_steps(ig_engine *h, ig_ensemble_id id, int64_t *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_ensemble_steps: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_steps: unknown ensemble id");
  return live_and_steps(h, E, nullptr, out);
}

static int build_flight(const ig_flight_params *p, FlightK &k) {
  std::memset(&k, 0, sizeof(k));
  if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return fail(IG_E_ARG, "ig_fly: dt must be > 0");
  if (p->rec_every < 1) return fail(IG_E_ARG, "ig_fly: rec_every must be >= 1");
  if (p->gas_model < IG_GAS_VACUUM || p->gas_model > IG_GAS_SDS)
    return fail(IG_E_ARG, "ig_fly: unknown gas model");
  if (p->gas_model == IG_GAS_HS && !(p->P_pa > 0.0 && p->T_k > 0.0 && p->c_star > 0.0))
    return fail(IG_E_ARG, "ig_fly: hard-sphere gas needs P_pa, T_k and c_star > 0");
  if (p->gas_model == IG_GAS_SDS && !p->sds_stats)
    return fail(IG_E_ARG, "ig_fly: SDS needs the (5, 1002) jump table");
  if (p->n_planes < 0 || p->n_planes > IG_MAX_PLANES)
    return fail(IG_E_ARG, "ig_fly: n_planes out of range");
  k.dt = p->dt;
  k.t_max = p->t_max;
  k.max_steps = p->max_steps;
  k.rec_every = p->rec_every;
  k.rec_rows = p->rec_rows;
  k.rec_mask = p->rec_mask;
  k.gas = p->gas_model;
  k.T_k = p->T_k;
  k.P_pa = p->P_pa;
  k.c_star = p->c_star;
  k.c_bar = p->c_bar;
  k.sig1d = p->sig1d;
  k.m_gas = p->m_gas;
  for (int i = 0; i < 3; ++i) k.vg[i] = p->flow[i];
  for (int i = 0; i < 6; ++i) {
    k.bnd_on[i] = p->bnd_on[i];
    k.bnd_val[i] = p->bnd_val[i];
  }
  k.n_pl = p->n_planes;
  for (int i = 0; i < p->n_planes; ++i) {
    if (p->pl_col[i] < 0 || p->pl_col[i] > 2) return fail(IG_E_ARG, "ig_fly: plane axis");
    k.pl_col[i] = p->pl_col[i];
    k.pl_val[i] = p->pl_val[i];
    k.pl_sgn[i] = p->pl_sgn[i];
    for (int c = 0; c < 4; ++c) k.pl_w[i][c] = p->pl_w[i][c];
    k.pl_kind[i] = p->pl_kind[i];
  }
  k.rng_key = p->rng_key;
  if (p->coll_batch < 0 || p->coll_batch > 64)
    return fail(IG_E_ARG, "ig_fly: coll_batch must be in [0, 64]");
  k.coll_batch = p->coll_batch; // 0: the precision's default, set per flight
  if (p->occupancy < 0 || p->occupancy > 4)
    return fail(IG_E_ARG, "ig_fly: occupancy must be 0 (default), 1, 2, 3 or 4");
  k.occupancy = p->occupancy;
  if (p->block != 0 && (p->block < 64 || p->block > 256 || p->block % 64 != 0))
    return fail(IG_E_ARG, "ig_fly: block must be 0 (default), 64, 128, 192 or 256");
  if (p->compact < 0 || p->compact > 3)
    return fail(IG_E_ARG, "ig_fly: compact must be 0, 1, 2 or 3");
  k.tp_col = -1;
  if (p->tp_axis != 0) {
    if (p->tp_axis < 1 || p->tp_axis > 3)
      return fail(IG_E_ARG, "ig_fly: transporter axis must be 0 (none), 1 (x), 2 (y) or 3 (z)");
    if (p->tp_accept_mm == p->tp_emit_mm)
      return fail(IG_E_ARG, "transporter: accept_mm == emit_mm -- a zero-length transporter "
                            "is a no-op in disguise");
    if (p->tp_max_passes < 1) return fail(IG_E_ARG, "ig_fly: transporter max_passes must be >= 1");
    if (p->gas_model == IG_GAS_SDS)
      return fail(IG_E_ARG, "ig_fly: the transporter belongs to the 3-D RK4 flight, not SDS");
    k.tp_col = p->tp_axis - 1;
    k.tp_accept = p->tp_accept_mm;
    k.tp_d = p->tp_emit_mm - p->tp_accept_mm;
    k.tp_sgn = p->tp_direction;
    k.tp_maxp = p->tp_max_passes;
  }
  return 0;
}

// The live ions of an ensemble as a dense index list: one ballot per
// wave, one atomic per wave for the wave's slice, lanes in order within
// it (waves in any order: an ion's flight does not depend on its lane).
__global__ void compact_live_kernel(const int32_t *__restrict__ fate, int64_t n,
                                    int32_t *__restrict__ idx, uint32_t *__restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n && fate[i] == IG_FATE_FLYING;
  const uint64_t m = __ballot(live);
  if (m == 0) return;
  const int lane = (int)(threadIdx.x & 63u);
  const int first = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, first);
  if (live) idx[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (int32_t)i;
}

__global__ void zero_u32_kernel(uint32_t *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

// Cell order of the live list (compact 2): a counting sort of the live
// ions by the grid cell they sit in, so that the lanes of a wave gather
// from the same or neighbouring field records.  The key is the linear
// cell index (axial, radial for r-z; x, y for planar; x, y, z for 3-D,
// the last fastest) modulo kMaxBlocks: exact cell order wherever the
// occupied cells span fewer than kMaxBlocks indices, and a few far-apart
// cells sharing a bucket elsewhere.  Within a cell the order is whatever
// the atomics give (an ion's flight does not depend on its lane).
constexpr int kMaxBlocks = 1 << 18;

struct BlockGrid {
  int32_t route, dims;
  int64_t n[3];       // nodes per axis (r-z: the half grid)
  double inv_h;       // node coordinates of the pack (1 / pitch)
};

static BlockGrid block_grid(const FieldsK &f) {
  BlockGrid g{};
  g.route = f.route;
  g.inv_h = 1.0 / f.h;
  g.n[0] = f.n0;
  g.n[1] = f.route == IG_ROUTE_RZ ? (f.n1 + 1) / 2 : f.n1;
  g.n[2] = f.route == IG_ROUTE_XYZ ? f.n2 : 1;
  g.dims = f.route == IG_ROUTE_XYZ ? 3 : 2;
  return g;
}

__device__ __forceinline__ uint32_t block_of(const BlockGrid &g, double x, double y, double z) {
  double c[3];
  c[0] = x * g.inv_h;
  if (g.route == IG_ROUTE_RZ) {
    c[1] = sqrt(y * y + z * z) * g.inv_h;
    c[2] = 0.0;
  } else {
    c[1] = y * g.inv_h;
    c[2] = g.dims == 3 ? z * g.inv_h : 0.0;
  }
  int64_t b[3];
  for (int a = 0; a < 3; ++a) {
    // outside the grid (or not a number): the nearest cell
    const double v = c[a] >= 0.0 ? c[a] : 0.0;
    const int64_t q = v < 4.0e18 ? (int64_t)v : g.n[a] - 1;
    b[a] = q < g.n[a] ? q : g.n[a] - 1;
  }
  const int64_t cell = (b[0] * g.n[1] + b[1]) * g.n[2] + b[2];
  return (uint32_t)(cell & (int64_t)(kMaxBlocks - 1));
}

__global__ void zero_blocks_kernel(uint32_t *__restrict__ hist, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) hist[i] = 0u;
}

__global__ void block_hist_kernel(const BlockGrid g, const EnsK e, const int32_t *__restrict__ idx,
                                  const uint32_t *__restrict__ count, uint32_t *__restrict__ key,
                                  uint32_t *__restrict__ hist) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)*count) return;
  const int32_t i = idx[t];
  const uint32_t k = block_of(g, e.x[i], e.y[i], e.z[i]);
  key[t] = k;
  atomicAdd(&hist[k], 1u);
}

// exclusive scan of the kMaxBlocks counts in two passes of kScanTiles
// workgroups of 256 lanes, each lane holding four consecutive counts (one
// 16-byte load): the tile sums, then each tile adds the sum of the tiles
// before it to its own scan
constexpr int kScanTile = 1024, kScanTiles = kMaxBlocks / kScanTile;
static_assert(kScanTiles <= 256, "one lane per earlier tile");

__device__ __forceinline__ uint32_t block_sum_256(uint32_t v, uint32_t *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[wv] = v;
  __syncthreads();
  const uint32_t tot = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return tot;
}

__global__ __launch_bounds__(256) void scan_sums_kernel(const uint32_t *__restrict__ hist,
                                                        uint32_t *__restrict__ sums) {
  __shared__ uint32_t red[4];
  const uint4 v = reinterpret_cast<const uint4 *>(hist + (size_t)blockIdx.x * kScanTile)[threadIdx.x];
  const uint32_t tot = block_sum_256(v.x + v.y + v.z + v.w, red);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void scan_apply_kernel(uint32_t *__restrict__ hist,
                                                         const uint32_t *__restrict__ sums) {
  __shared__ uint32_t red[4], wsum[4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t base = block_sum_256(tid < (int)blockIdx.x ? sums[tid] : 0u, red);
  uint4 *p = reinterpret_cast<uint4 *>(hist + (size_t)blockIdx.x * kScanTile) + tid;
  const uint4 v = *p;
  const uint32_t own = v.x + v.y + v.z + v.w;
  // inclusive scan of the lanes' totals within the wave, then across waves
  uint32_t inc = own;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  uint32_t run = base + inc - own;
  for (int k = 0; k < wv; ++k) run += wsum[k];
  uint4 o;
  o.x = run;
  o.y = o.x + v.x;
  o.z = o.y + v.y;
  o.w = o.z + v.z;
  *p = o;
}

__global__ void block_scatter_kernel(const int32_t *__restrict__ idx,
                                     const uint32_t *__restrict__ count,
                                     const uint32_t *__restrict__ key,
                                     uint32_t *__restrict__ offsets, int32_t *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)*count) return;
  out[atomicAdd(&offsets[key[t]], 1u)] = idx[t];
}

// the compaction buffers, sized for the ensemble's capacity (allocated
// once, before any capture)
static int live_bufs(Ensemble *E, bool ordered) {
  const size_t cap = (size_t)std::max<int64_t>(E->cap, 1);
  if (!E->live_idx.p)
    if (int rc = dev_alloc(E->live_idx, cap * sizeof(int32_t))) return rc;
  if (!E->live_cnt.p)
    if (int rc = dev_alloc(E->live_cnt, sizeof(uint32_t))) return rc;
  if (ordered) {
    if (!E->live_key.p)
      if (int rc = dev_alloc(E->live_key, cap * sizeof(uint32_t))) return rc;
    if (!E->live_sorted.p)
      if (int rc = dev_alloc(E->live_sorted, cap * sizeof(int32_t))) return rc;
    if (!E->live_hist.p)
      if (int rc = dev_alloc(E->live_hist, (size_t)(kMaxBlocks + kScanTiles) * sizeof(uint32_t)))
        return rc;
  }
  return 0;
}

// the kernel arguments of one window: with compaction, the live list is
// rebuilt on the stream first, and with compact 2 put in cell order
// (kernels only, so a graph may capture it)
// the live-list mode of a flight: compact 3 lets the engine choose per
// route -- cell order for the 3-D route, whose field gathers reach HBM and
// gain most from neighbouring lanes sharing records (1.8x on the SLIM
// board), slot order elsewhere, where the pack is cache-resident and the
// lanes' coalesced state loads matter more (cell order costs 3 % on the
// funnel)
static int compact_mode(const ig_flight_params *p, const FieldsK &f) {
  if (p->compact != 3) return p->compact;
  return f.route == IG_ROUTE_XYZ ? 2 : 1;
}

// where a window's live list is built: the ensemble's own buffers
// (ig_fly, ig_fly_graph) or a device's no-allocation scratch
struct LiveBufs {
  int32_t *idx = nullptr, *sorted = nullptr;
  uint32_t *cnt = nullptr, *key = nullptr, *hist = nullptr;
};

static_assert((size_t)(kMaxBlocks + kScanTiles) <= kNaHistWords,
              "the no-allocation scratch holds the cell histogram");

static int window_ens(ig_engine *h, Ensemble *E, const FieldsK &f, const ig_flight_params *p,
                      EnsK &ek, const LiveBufs *fixed = nullptr) {
  ek = E->k;
  const int mode = compact_mode(p, f);
  if (!mode || E->n == 0) return 0;
  const bool ordered = mode == 2;
  LiveBufs b;
  if (fixed) {
    b = *fixed;
  } else {
    if (int rc = live_bufs(E, ordered)) return rc;
    b.idx = (int32_t *)E->live_idx.p;
    b.cnt = (uint32_t *)E->live_cnt.p;
    b.key = (uint32_t *)E->live_key.p;
    b.sorted = (int32_t *)E->live_sorted.p;
    b.hist = (uint32_t *)E->live_hist.p;
  }
  const unsigned grid = (unsigned)((E->n + 255) / 256);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, h->stream, b.cnt);
  hipLaunchKernelGGL(compact_live_kernel, dim3(grid), dim3(256), 0, h->stream, E->k.fate, E->n,
                     b.idx, b.cnt);
  ek.idx = (const int32_t *)b.idx;
  ek.n_live = (const uint32_t *)b.cnt;
  if (ordered) {
    const BlockGrid g = block_grid(f);
    const int64_t nb = kMaxBlocks;
    hipLaunchKernelGGL(zero_blocks_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                       h->stream, b.hist, nb);
    hipLaunchKernelGGL(block_hist_kernel, dim3(grid), dim3(256), 0, h->stream, g, E->k,
                       (const int32_t *)b.idx, (const uint32_t *)b.cnt, b.key, b.hist);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(kScanTiles), dim3(256), 0, h->stream,
                       (const uint32_t *)b.hist, b.hist + kMaxBlocks);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(kScanTiles), dim3(256), 0, h->stream, b.hist,
                       (const uint32_t *)(b.hist + kMaxBlocks));
    hipLaunchKernelGGL(block_scatter_kernel, dim3(grid), dim3(256), 0, h->stream,
                       (const int32_t *)b.idx, (const uint32_t *)b.cnt,
                       (const uint32_t *)b.key, b.hist, b.sorted);
    ek.idx = (const int32_t *)b.sorted;
  }
  return 0;
}

// Validation and device-side setup shared by ig_fly and ig_fly_graph:
// the kernel arguments for one window of (fid, p, eid).
static int prepare_flight(ig_engine *h, ig_fields_id fid, const ig_flight_params *p,
                          ig_ensemble_id eid, const char *who, const FieldsK *&fk,
                          FlightK &k, Ensemble *&E, NoAllocDevice *na = nullptr) {
  const std::string w(who);
  if (!h || !p) return fail(IG_E_ARG, w + ": null argument");
  auto fit = h->fields.find(fid);
  if (fit == h->fields.end()) return fail(IG_E_STATE, w + ": unknown field id");
  E = find_ens(h, eid);
  if (!E) return fail(IG_E_STATE, w + ": unknown ensemble id");
  if (int rc = build_flight(p, k)) return rc;
  if (p->rec_rows > 0 && (p->rec_rows != E->rec_rows || E->k.rec == nullptr))
    return fail(IG_E_ARG, w + ": rec_rows differs from the ensemble's record buffer");
  k.rec_cols = E->rec_cols;
  if (int rc = set_dev(h)) return rc;
  if (k.gas == IG_GAS_SDS) {
    // the (5, 1002) jump table stays on the device until it changes: in
    // the engine's own buffer, or without allocation in the device's
    // scratch
    const size_t nt = kNaSdsDoubles;
    std::vector<double> &host = na ? na->sds_host : h->sds_host;
    if (host.size() != nt ||
        std::memcmp(host.data(), p->sds_stats, nt * sizeof(double)) != 0) {
      HIPX(hipStreamSynchronize(h->stream));
      host.assign(p->sds_stats, p->sds_stats + nt);
      void *dst = na ? (void *)(na->ptr + kNaSds) : h->sds.p;
      if (!dst) {
        if (int rc = dev_alloc(h->sds, nt * sizeof(double))) return rc;
        dst = h->sds.p;
      }
      HIPX(hipMemcpyAsync(dst, host.data(), nt * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
      HIPX(hipStreamSynchronize(h->stream));
    }
    k.sds_stats = na ? (const double *)(na->ptr + kNaSds) : (const double *)h->sds.p;
  }
  fk = &fit->second->k;
  if (int rc = bind_tunes(h, E, fid, *fit->second, w)) return rc;
  const int32_t clk = (fk->route == IG_ROUTE_XYZ && k.gas != IG_GAS_SDS) ? 1 : 0;
  if (E->clock >= 0 && (E->clock != clk || E->route != fk->route))
    return fail(IG_E_ARG, w + ": the ensemble was flown on another route or clock (3-D RK4 "
                              "carries s and m/s, every other flight us and mm/us); load it "
                              "again");
  if (k.tp_col >= 0 && fk->route != IG_ROUTE_XYZ)
    return fail(IG_E_ARG, w + ": the transporter belongs to the 3-D route");
  if (k.tp_col >= 0 && p->rec_rows > 0 && E->rec_cols < 11)
    return fail(IG_E_ARG, w + ": records with the transporter need 11 columns (the pass count)");
  E->route = fk->route;
  E->clock = clk;
  if (k.occupancy == 0) k.occupancy = default_occupancy(fk->prec, fk->route, k.gas);
  if (k.coll_batch == 0)
    k.coll_batch = fk->prec != IG_PREC_F32 ? kDefaultCollBatch
                   : (fk->route == IG_ROUTE_XYZ ? kDefaultCollBatchFlat3D : kDefaultCollBatchFlat);
  return 0;
}

static int launch_error(int rc, const char *who) {
  if (rc == -2) return fail(IG_E_ARG, std::string(who) + ": no kernel instantiated for this field pack");
  hipError_t e = hipGetLastError();
  return fail(IG_E_HIP, std::string(who) + ": launch failed: " + hipGetErrorString(e));
}

int ig_fly(ig_engine *h, ig_fields_id fid, const ig_flight_params *p, ig_ensemble_id eid) {
  const FieldsK *f = nullptr;
  FlightK k;
  Ensemble *E = nullptr;
  if (int rc = prepare_flight(h, fid, p, eid, "ig_fly", f, k, E)) return rc;
  HIPX(hipEventRecord(h->ev0, h->stream));
  EnsK ek;
  if (int rc = window_ens(h, E, *f, p, ek)) return rc;
  int rc = launch_route(*f, k, ek, E->n, launch_block(h, p, E->n), h->stream);
  HIPX(hipEventRecord(h->ev1, h->stream));
  h->timed = true;
  if (rc) return launch_error(rc, "ig_fly");
  return 0;
}

int ig_fly_noalloc(ig_engine *h, ig_fields_id fid, const ig_flight_params *p,
                   ig_ensemble_id eid) {
  if (!h || !p) return fail(IG_E_ARG, "ig_fly_noalloc: null argument");
  NoAllocDevice *na = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    auto it = g_na.find(h->device);
    if (it == g_na.end()) return fail(IG_E_STATE, "ig_fly_noalloc: no scratch set up for device");
    if (it->second.status) return fail(it->second.status, "ig_fly_noalloc: " + it->second.err);
    na = &it->second;
  }
  Ensemble *E0 = find_ens(h, eid);
  if (E0 && E0->n > kNoAllocMaxIons)
    return fail(IG_E_ARG, "ig_fly_noalloc: " + std::to_string(E0->n) +
                              " ions exceed the no-allocation limit of " +
                              std::to_string(kNoAllocMaxIons) +
                              " (IG_NOALLOC_MAX_IONS); fly them with ig_fly");
  const FieldsK *f = nullptr;
  FlightK k;
  Ensemble *E = nullptr;
  if (int rc = prepare_flight(h, fid, p, eid, "ig_fly_noalloc", f, k, E, na)) return rc;
  LiveBufs b;
  b.cnt = (uint32_t *)na->ptr;
  b.idx = (int32_t *)(na->ptr + kNaIdx);
  b.key = (uint32_t *)(na->ptr + kNaKey);
  b.sorted = (int32_t *)(na->ptr + kNaSorted);
  b.hist = (uint32_t *)(na->ptr + kNaHist);
  HIPX(hipEventRecord(h->ev0, h->stream));
  EnsK ek;
  if (int rc = window_ens(h, E, *f, p, ek, &b)) return rc;
  int rc = launch_route(*f, k, ek, E->n, launch_block(h, p, E->n), h->stream);
  HIPX(hipEventRecord(h->ev1, h->stream));
  h->timed = true;
  if (rc) return launch_error(rc, "ig_fly_noalloc");
  return 0;
}

int ig_fly_graph(ig_engine *h, ig_fields_id fid, const ig_flight_params *p, ig_ensemble_id eid,
                 int32_t n_windows) {
  if (n_windows < 1 || n_windows > 4096)
    return fail(IG_E_ARG, "ig_fly_graph: n_windows must be in [1, 4096]");
  const FieldsK *f = nullptr;
  FlightK k;
  Ensemble *E = nullptr;
  if (int rc = prepare_flight(h, fid, p, eid, "ig_fly_graph", f, k, E)) return rc;
  const int block = launch_block(h, p, E->n);
  if (compact_mode(p, *f))
    if (int rc = live_bufs(E, compact_mode(p, *f) == 2)) return rc;
  // the graph is a function of the exact kernel arguments and launch shape
  std::vector<uint8_t> key(sizeof(FieldsK) + sizeof(FlightK) + sizeof(EnsK) + 3 * 8);
  {
    uint8_t *q = key.data();
    std::memcpy(q, f, sizeof(FieldsK));
    q += sizeof(FieldsK);
    std::memcpy(q, &k, sizeof(FlightK));
    q += sizeof(FlightK);
    std::memcpy(q, &E->k, sizeof(EnsK));
    q += sizeof(EnsK);
    const int64_t tail[3] = {E->n, block, n_windows + ((int64_t)p->compact << 40)};
    std::memcpy(q, tail, sizeof(tail));
  }
  hipGraphExec_t exec = nullptr;
  for (auto &g : h->graphs)
    if (g.key == key) exec = g.exec;
  if (!exec) {
    hipGraph_t graph = nullptr;
    HIPX(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int32_t wdx = 0; wdx < n_windows && rc == 0; ++wdx) {
      EnsK ek;
      rc = window_ens(h, E, *f, p, ek);
      if (rc == 0) rc = launch_route(*f, k, ek, E->n, block, h->stream);
    }
    hipError_t ce = hipStreamEndCapture(h->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return launch_error(rc, "ig_fly_graph");
    }
    if (ce != hipSuccess)
      return fail(IG_E_HIP, std::string("ig_fly_graph: capture failed: ") + hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess)
      return fail(IG_E_HIP, std::string("ig_fly_graph: instantiate failed: ") +
                                hipGetErrorString(ie));
    if (h->graphs.size() >= kMaxGraphs) {
      (void)hipGraphExecDestroy(h->graphs.front().exec);
      h->graphs.erase(h->graphs.begin());
    }
    h->graphs.push_back(GraphEntry{std::move(key), exec});
  }
  HIPX(hipEventRecord(h->ev0, h->stream));
  HIPX(hipGraphLaunch(exec, h->stream));
  HIPX(hipEventRecord(h->ev1, h->stream));
  h->timed = true;
  return 0;
}

// np.gradient's spacing for one axis from its coordinate array (numpy's
// own arithmetic: steps by np.diff, the scalar case when every step is
// equal, else the three-point coefficients a, b, c)
static void numpy_axis(const double *x, int64_t n, std::vector<double> &abc, GradAxis &g) {
  std::vector<double> d((size_t)(n - 1));
  for (int64_t i = 0; i + 1 < n; ++i) d[i] = x[i + 1] - x[i];
  bool uni = true;
  for (double v : d) uni = uni && v == d[0];
  g.n = n;
  g.uniform = uni ? 1 : 0;
  g.dx = d[0];
  g.dx_lo = d[0];
  g.dx_hi = d[n - 2];
  g.h_one = x[1] - x[0];
  abc.clear();
  if (!uni) {
    abc.resize((size_t)(3 * (n - 2)));
    for (int64_t i = 0; i + 2 < n; ++i) {
      const double dx1 = d[i], dx2 = d[i + 1];
      abc[3 * i] = -(dx2) / (dx1 * (dx1 + dx2));
      abc[3 * i + 1] = (dx2 - dx1) / (dx1 * dx2);
      abc[3 * i + 2] = dx1 / (dx2 * (dx1 + dx2));
    }
  }
}

int ig_bases_upload(ig_engine *h, const ig_bases_desc *d, ig_bases_id *out) {
  if (!h || !d || !out || !d->phi || !d->metal)
    return fail(IG_E_ARG, "ig_bases_upload: null argument");
  const bool rz = d->route == IG_ROUTE_RZ;
  if (!rz && d->route != IG_ROUTE_XYZ) return fail(IG_E_ARG, "ig_bases_upload: unknown route");
  if (d->n0 < 2 || d->n1 < 2 || (rz ? d->n2 != 1 : d->n2 < 2) || d->n_bases < 1)
    return fail(IG_E_ARG, "ig_bases_upload: bad grid or base count");
  if (!(d->base_volts > 0.0)) return fail(IG_E_ARG, "ig_bases_upload: base_volts must be > 0");
  if (rz && (!d->coord0 || !d->coord1))
    return fail(IG_E_ARG, "ig_bases_upload: r-z bases need coord0 and coord1");
  if (!rz && !(d->h_mm > 0.0)) return fail(IG_E_ARG, "ig_bases_upload: h_mm must be > 0");
  if (int rc = set_dev(h)) return rc;
  auto B = std::make_unique<Bases>();
  B->route = d->route;
  B->n0 = d->n0;
  B->n1 = d->n1;
  B->n2 = d->n2;
  B->n_bases = d->n_bases;
  B->base_volts = d->base_volts;
  const int64_t nodes = d->n0 * d->n1 * d->n2;
  int rc = upload(h, B->phi, d->phi, (size_t)(nodes * d->n_bases) * sizeof(double));
  if (rc) return rc;
  std::vector<uint8_t> m((size_t)nodes);
  for (int64_t n = 0; n < nodes; ++n) m[n] = d->metal[n] ? 1 : 0;
  if ((rc = upload(h, B->metal, m.data(), m.size()))) return rc;
  if ((rc = dev_alloc(B->work, (size_t)nodes * sizeof(double)))) return rc;
  if (rz) {
    std::vector<double> a0, a1;
    numpy_axis(d->coord0, d->n0, a0, B->g0);
    numpy_axis(d->coord1, 2 * d->n1 - 1, a1, B->g1);
    if (!a0.empty() && (rc = upload(h, B->abc0, a0.data(), a0.size() * sizeof(double)))) return rc;
    if (!a1.empty() && (rc = upload(h, B->abc1, a1.data(), a1.size() * sizeof(double)))) return rc;
    B->g0.abc = (const double *)B->abc0.p;
    B->g1.abc = (const double *)B->abc1.p;
  } else {
    // np.gradient(phi, h_mm): one scalar spacing on every axis
    B->g0 = GradAxis{0, 1, d->h_mm, nullptr, d->h_mm, d->h_mm, d->h_mm};
  }
  HIPX(hipStreamSynchronize(h->stream));
  const int64_t id = h->next_id++;
  h->bases[id] = std::move(B);
  *out = id;
  return 0;
}

int ig_bases_release(ig_engine *h, ig_bases_id id) {
  if (!h) return fail(IG_E_ARG, "ig_bases_release: null engine");
  auto it = h->bases.find(id);
  if (it == h->bases.end()) return fail(IG_E_STATE, "ig_bases_release: unknown bases id");
  if (int rc = set_dev(h)) return rc;
  HIPX(hipStreamSynchronize(h->stream));
  h->bases.erase(it);
  return 0;
}

int ig_fields_compose(ig_engine *h, ig_bases_id bid, ig_fields_id fid, int32_t set,
                      int32_t n_terms, const int32_t *base_index, const double *weight) {
  if (!h || (n_terms > 0 && (!base_index || !weight)))
    return fail(IG_E_ARG, "ig_fields_compose: null argument");
  auto fit = h->fields.find(fid);
  if (fit == h->fields.end()) return fail(IG_E_STATE, "ig_fields_compose: unknown field id");
  Fields &F = *fit->second;
  if (F.fold.on)
    return fail(IG_E_ARG, "ig_fields_compose: the pack's linked drive channels were folded at "
                          "upload (IG_FOLD_LINKED); upload the pack to compose into with "
                          "IG_FOLD_NONE");
  auto bit = h->bases.find(bid);
  if (bit == h->bases.end()) return fail(IG_E_STATE, "ig_fields_compose: unknown bases id");
  Bases &B = *bit->second;
  const FieldsK &k = F.k;
  const bool rz = B.route == IG_ROUTE_RZ;
  if (k.route != B.route) return fail(IG_E_ARG, "ig_fields_compose: route differs from the pack's");
  if (rz ? (k.n0 != B.n0 || k.n1 != 2 * B.n1 - 1)
         : (k.n0 != B.n0 || k.n1 != B.n1 || k.n2 != B.n2))
    return fail(IG_E_ARG, "ig_fields_compose: bases grid does not match the pack");
  if (set < 0 || set >= k.n_sets) return fail(IG_E_ARG, "ig_fields_compose: set out of range");
  if (n_terms < 0 || n_terms > (1 << 20)) return fail(IG_E_ARG, "ig_fields_compose: bad term count");
  for (int32_t t = 0; t < n_terms; ++t)
    if (base_index[t] < 0 || base_index[t] >= B.n_bases)
      return fail(IG_E_ARG, "ig_fields_compose: base index out of range");
  if (int rc = set_dev(h)) return rc;
  const size_t nt = (size_t)std::max<int32_t>(n_terms, 1);
  if (B.terms_i.bytes < nt * sizeof(int32_t)) {
    HIPX(hipStreamSynchronize(h->stream));
    B.terms_i.reset();
    B.terms_w.reset();
    if (int rc = dev_alloc(B.terms_i, nt * sizeof(int32_t))) return rc;
    if (int rc = dev_alloc(B.terms_w, nt * sizeof(double))) return rc;
  }
  if (n_terms > 0) {
    HIPX(hipMemcpyAsync(B.terms_i.p, base_index, n_terms * sizeof(int32_t),
                        hipMemcpyHostToDevice, h->stream));
    HIPX(hipMemcpyAsync(B.terms_w.p, weight, n_terms * sizeof(double), hipMemcpyHostToDevice,
                        h->stream));
  }
  const int64_t nodes = B.n0 * B.n1 * B.n2;
  hipLaunchKernelGGL(compose_kernel, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0,
                     h->stream, (const double *)B.phi.p, nodes, n_terms,
                     (const int32_t *)B.terms_i.p, (const double *)B.terms_w.p, B.base_volts,
                     (double *)B.work.p);
  const double inv = 1.0 / k.scale[set];
  const bool f64 = k.prec == IG_PREC_F64;
  if (rz) {
    const int64_t out_nodes = B.n0 * (2 * B.n1 - 1);
    dim3 g((unsigned)((out_nodes + 255) / 256));
    if (f64)
      hipLaunchKernelGGL(grad_rz_kernel<double>, g, dim3(256), 0, h->stream,
                         (const double *)B.work.p, (const uint8_t *)B.metal.p, B.n0, B.n1, B.g0,
                         B.g1, F.grid.p, (int64_t)k.stride, set, inv);
    else
      hipLaunchKernelGGL(grad_rz_kernel<float>, g, dim3(256), 0, h->stream,
                         (const double *)B.work.p, (const uint8_t *)B.metal.p, B.n0, B.n1, B.g0,
                         B.g1, F.grid.p, (int64_t)k.stride, set, inv);
  } else {
    dim3 g((unsigned)((nodes + 255) / 256));
    if (f64)
      hipLaunchKernelGGL(grad_3d_kernel<double>, g, dim3(256), 0, h->stream,
                         (const double *)B.work.p, (const uint8_t *)B.metal.p, B.n0, B.n1, B.n2,
                         B.g0, F.grid.p, (int64_t)k.stride, set, inv);
    else
      hipLaunchKernelGGL(grad_3d_kernel<float>, g, dim3(256), 0, h->stream,
                         (const double *)B.work.p, (const uint8_t *)B.metal.p, B.n0, B.n1, B.n2,
                         B.g0, F.grid.p, (int64_t)k.stride, set, inv);
  }
  HIPX(hipGetLastError());
  // the host term arrays are the caller's: finish the copies before return
  HIPX(hipStreamSynchronize(h->stream));
  return 0;
}

int ig_fields_layout(ig_engine *h, ig_fields_id fid, int32_t *n_sets, int32_t *n_chan,
                     int32_t *stride, int32_t *n_chan_given) {
  if (!h || !n_sets || !n_chan || !stride || !n_chan_given)
    return fail(IG_E_ARG, "ig_fields_layout: null argument");
  auto fit = h->fields.find(fid);
  if (fit == h->fields.end()) return fail(IG_E_STATE, "ig_fields_layout: unknown field id");
  const Fields &F = *fit->second;
  *n_sets = F.k.n_sets;
  *n_chan = F.k.K;
  *stride = F.k.stride;
  *n_chan_given = F.fold.on ? F.fold.K0 : F.k.K;
  return 0;
}

int ig_fields_read_set(ig_engine *h, ig_fields_id fid, int32_t set, int32_t comp, double *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_fields_read_set: null argument");
  auto fit = h->fields.find(fid);
  if (fit == h->fields.end()) return fail(IG_E_STATE, "ig_fields_read_set: unknown field id");
  Fields &F = *fit->second;
  const FieldsK &k = F.k;
  if (set < 0 || set >= k.n_sets || comp < 0 || comp >= k.n_comp)
    return fail(IG_E_ARG, "ig_fields_read_set: set or component out of range");
  if (int rc = set_dev(h)) return rc;
  const int64_t nodes = k.n0 * k.n1 * k.n2;
  DevBuf tmp;
  if (int rc = dev_alloc(tmp, (size_t)nodes * sizeof(double))) return rc;
  const int32_t q = set * k.n_comp + comp;
  dim3 g((unsigned)((nodes + 255) / 256));
  if (k.prec == IG_PREC_F64)
    hipLaunchKernelGGL(read_comp_kernel<double>, g, dim3(256), 0, h->stream, F.grid.p, nodes,
                       (int64_t)k.stride, q, k.scale[set], (double *)tmp.p);
  else
    hipLaunchKernelGGL(read_comp_kernel<float>, g, dim3(256), 0, h->stream, F.grid.p, nodes,
                       (int64_t)k.stride, q, k.scale[set], (double *)tmp.p);
  HIPX(hipGetLastError());
  HIPX(hipMemcpyAsync(out, tmp.p, (size_t)nodes * sizeof(double), hipMemcpyDeviceToHost,
                      h->stream));
  HIPX(hipStreamSynchronize(h->stream));
  return 0;
}

// a scratch buffer of at least `bytes` (grown, never shrunk)
static int grow_buf(DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return 0;
  b.reset();
  return dev_alloc(b, bytes);
}

// Births of a source (ig_births.h): validate, size the word stream and the
// jump table, run the passes, and grow the stream when an ion would start
// past it.  `out` is filled for the destination; `n_out` receives N.
static int births_run(ig_engine *h, const ig_births_desc *d, const char *who,
                      const std::function<BirthsOut(int64_t)> &dest, int64_t *n_out) {
  const std::string w(who);
  if (!d) return fail(IG_E_ARG, w + ": null source");
  if (d->distribution < IG_BIRTH_POINT || d->distribution > IG_BIRTH_GAUSSIAN)
    return fail(IG_E_ARG, w + ": unknown distribution " + std::to_string(d->distribution));
  if (d->n_per < 1 || d->n_mz < 1 || !d->mz_list)
    return fail(IG_E_ARG, w + ": n_per >= 1 ions per mass and n_mz >= 1 masses are required");
  if (d->axis < 0 || d->axis > 2) return fail(IG_E_ARG, w + ": axis must be 0, 1 or 2");
  if (!d->zig_k || !d->zig_w || !d->zig_f)
    return fail(IG_E_ARG, w + ": the standard-normal layer tables are required");
  if (d->distribution == IG_BIRTH_GAUSSIAN && d->gauss_tries < 1)
    return fail(IG_E_ARG, w + ": gauss_tries must be >= 1");
  if ((d->pcg_inc[1] & 1u) == 0) return fail(IG_E_ARG, w + ": the PCG64 increment must be odd");
  const int64_t n = d->n_per * (int64_t)d->n_mz;
  *n_out = n;
  if (int rc = set_dev(h)) return rc;
  // expected words per ion: one per uniform, ~1.03 per normal, 1/p per
  // truncated-Gaussian axis with p the probability of landing inside
  double est = 1.0; // the birth time
  if (d->distribution == IG_BIRTH_DISC) est += 2.0;
  if (d->distribution == IG_BIRTH_BOX) est += 3.0;
  if (d->distribution == IG_BIRTH_GAUSSIAN)
    for (int k = 0; k < 3; ++k)
      if (d->gauss_sigma[k] > 0.0) {
        const double acc = std::erf(d->gauss_trunc[k] / (d->gauss_sigma[k] * std::sqrt(2.0)));
        est += 1.03 / std::max(acc, 1e-3);
      }
  if (d->has_ke) est += 1.0;
  if (d->temperature_k > 0.0) est += 3.09;
  if (d->has_dv) est += 3.09;
  BirthsK b{};
  b.dist = d->distribution;
  b.axis = d->axis;
  b.n_per = d->n_per;
  b.n = n;
  b.n_mz = d->n_mz;
  b.grid_g = (int64_t)std::ceil(std::sqrt((double)d->n_per));
  for (int k = 0; k < 3; ++k) {
    b.p0[k] = d->p0[k];
    b.box[k] = d->box_mm[k];
    b.gsig[k] = d->gauss_sigma[k];
    b.gtrunc[k] = d->gauss_trunc[k];
    b.dir[k] = d->dir_unit[k];
    b.dvs[k] = d->dv_sigma[k];
    b.vd[k] = d->v_drift[k];
  }
  b.r_mm = d->r_mm;
  b.len_mm = d->len_mm;
  b.gtries = d->gauss_tries;
  b.has_ke = d->has_ke;
  b.has_dir = d->has_dir;
  b.has_dv = d->has_dv;
  b.has_drift = d->has_drift;
  b.ke_lo = d->ke_lo;
  b.ke_hi = d->ke_hi;
  b.temp_k = d->temperature_k;
  b.tob_span = d->tob_span_us;
  // tables: 256 thresholds, 256 widths, 256 densities, the masses
  const size_t tab_bytes = 256 * 8 * 3 + (size_t)d->n_mz * 8;
  if (int rc = grow_buf(h->btab, tab_bytes)) return rc;
  uint8_t *tb = (uint8_t *)h->btab.p;
  HIPX(hipMemcpyAsync(tb, d->zig_k, 256 * 8, hipMemcpyHostToDevice, h->stream));
  HIPX(hipMemcpyAsync(tb + 2048, d->zig_w, 256 * 8, hipMemcpyHostToDevice, h->stream));
  HIPX(hipMemcpyAsync(tb + 4096, d->zig_f, 256 * 8, hipMemcpyHostToDevice, h->stream));
  HIPX(hipMemcpyAsync(tb + 6144, d->mz_list, (size_t)d->n_mz * 8, hipMemcpyHostToDevice,
                      h->stream));
  b.zk = (const uint64_t *)tb;
  b.zw = (const double *)(tb + 2048);
  b.zf = (const double *)(tb + 4096);
  b.mz = (const double *)(tb + 6144);
  int32_t levels = 1;
  while (levels < 62 && ((n - 1) >> levels) > 0) ++levels;
  if (int rc = grow_buf(h->bflag, 16)) return rc;
  const BirthsOut out = dest(n);
  for (int c = 0; c < 7; ++c)
    if (!out.col[c]) return fail(IG_E_NOMEM, w + ": no device memory for the births");
  double grow = 1.15;
  for (int attempt = 0; attempt < 8; ++attempt, grow *= 1.5) {
    const int64_t P = (int64_t)std::ceil((double)n * est * grow) + 4096;
    const int64_t nW = P + 65536;
    if (P + 2 > (int64_t)INT32_MAX / 2)
      return fail(IG_E_ARG, w + ": " + std::to_string(n) + " births need more stream positions "
                                "than one call indexes; split the source");
    if (int rc = grow_buf(h->bw, (size_t)nW * 8)) return rc;
    if (int rc = grow_buf(h->bj, (size_t)levels * (size_t)(P + 2) * 4)) return rc;
    if (births_launch(b, d->pcg_state[0], d->pcg_state[1], d->pcg_inc[0], d->pcg_inc[1],
                      (uint64_t *)h->bw.p, nW, (int32_t *)h->bj.p, P, levels, out,
                      (int64_t *)h->bflag.p, h->stream))
      return fail(IG_E_HIP, w + ": births launch failed: " + hipGetErrorString(hipGetLastError()));
    int64_t flag[2] = {0, 0};
    HIPX(hipMemcpyAsync(flag, h->bflag.p, 16, hipMemcpyDeviceToHost, h->stream));
    HIPX(hipStreamSynchronize(h->stream));
    if (flag[0] & 1) continue; // an ion starts past the stream: grow it
    if (flag[0] & 2)
      return fail(IG_E_BIRTH_TRIES,
                  w + ": ion " + std::to_string(flag[1]) +
                      ": no draw of the truncated Gaussian fell inside the truncation in " +
                      std::to_string(d->gauss_tries) + " tries");
    return 0;
  }
  return fail(IG_E_STATE, w + ": the word stream could not be sized for the births");
}

int ig_births(ig_engine *h, const ig_births_desc *d, double *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_births: null argument");
  int64_t n = 0;
  int rc = births_run(h, d, "ig_births",
                      [&](int64_t nn) {
                        BirthsOut o{};
                        if (grow_buf(h->bout, (size_t)std::max<int64_t>(nn, 1) * 7 * 8) == 0)
                          for (int c = 0; c < 7; ++c) o.col[c] = (double *)h->bout.p + c;
                        o.stride = 7;
                        return o;
                      },
                      &n);
  if (rc) return rc;
  HIPX(hipMemcpyAsync(out, h->bout.p, (size_t)n * 7 * 8, hipMemcpyDeviceToHost, h->stream));
  HIPX(hipStreamSynchronize(h->stream));
  return 0;
}

int ig_ensemble_births_at(ig_engine *h, ig_ensemble_id id, int64_t offset,
                          const ig_births_desc *d) {
  if (!h || !d) return fail(IG_E_ARG, "ig_ensemble_births_at: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_births_at: unknown ensemble id");
  const int64_t n = d->n_per * (int64_t)d->n_mz;
  if (offset < 0 || n < 0 || offset + n > E->n)
    return fail(IG_E_ARG, "ig_ensemble_births_at: slots [" + std::to_string(offset) + ", " +
                              std::to_string(offset + n) + ") are not all loaded");
  int64_t nn = 0;
  const EnsK &k = E->k;
  return births_run(h, d, "ig_ensemble_births_at",
                    [&](int64_t) {
                      BirthsOut o{};
                      double *cols[7] = {k.x, k.y, k.z, k.vx, k.vy, k.vz, k.tob};
                      for (int c = 0; c < 7; ++c) o.col[c] = cols[c] + offset;
                      o.stride = 1;
                      return o;
                    },
                    &nn);
}

// the birth refusals of the three routes (ig_ensemble_refuse_births)
__global__ void refuse_births_kernel(const FieldsK f, const EnsK e, int64_t off, int64_t n,
                                     unsigned long long *count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t i = off + t;
  if (e.fate[i] != IG_FATE_FLYING || e.steps[i] != 0) return;
  const double tol = 1e-9; // birth_domain.DOMAIN_EDGE_TOL
  const double x = e.x[i], y = e.y[i], z = e.z[i];
  int fate = -1, label = 0;
  auto inside = [&](double g, int64_t nn) { return -tol <= g && g <= (double)(nn - 1) + tol; };
  if (f.route == IG_ROUTE_XYZ) {
    const double inv_h = 1.0 / f.h;
    const double gx = x * inv_h, gy = y * inv_h, gz = z * inv_h;
    if (!inside(gx, f.n0) || !inside(gy, f.n1) || !inside(gz, f.n2))
      fate = IG_FATE_BOX;
    else if (in_metal_3d(f.metal, gx, gy, gz, f.n0, f.n1, f.n2))
      fate = IG_FATE_IMPACT;
  } else {
    // r-z: the half grid (n0, nr) of the unfolded pack (n0, 2 nr - 1);
    // planar: the pack's own (n0, n1)
    const bool rz = f.route == IG_ROUTE_RZ;
    const int64_t ncol = f.n1, nr = rz ? (f.n1 + 1) / 2 : f.n1, c0 = rz ? nr - 1 : 0;
    const double g0 = x / f.h, g1 = (rz ? hypot(y, z) : y) / f.h;
    if (!inside(g0, f.n0) || !inside(g1, nr)) {
      fate = IG_FATE_BOX;
    } else {
      const double fi = floor(g0 + 0.5), fj = floor(g1 + 0.5);
      if (fi >= 0.0 && fj >= 0.0 && fi <= (double)(f.n0 - 1) && fj <= (double)(nr - 1)) {
        const int64_t node = (int64_t)fi * ncol + c0 + (int64_t)fj;
        if (f.metal[node]) {
          fate = IG_FATE_IMPACT;
          label = f.label ? (int)f.label[node] : 0;
        }
      }
    }
  }
  if (fate < 0) return;
  e.fate[i] = fate;
  e.face[i] = IG_FACE_BIRTH;
  e.label[i] = label;
  atomicAdd(count, 1ull);
}

int ig_ensemble_refuse_births(ig_engine *h, ig_ensemble_id id, ig_fields_id fid, int64_t offset,
                              int64_t n, int64_t *n_refused) {
  if (!h) return fail(IG_E_ARG, "ig_ensemble_refuse_births: null engine");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_refuse_births: unknown ensemble id");
  auto fit = h->fields.find(fid);
  if (fit == h->fields.end())
    return fail(IG_E_STATE, "ig_ensemble_refuse_births: unknown field id");
  const FieldsK &fk = fit->second->k;
  if (offset < 0 || n < 0 || offset + n > E->n)
    return fail(IG_E_ARG, "ig_ensemble_refuse_births: slots [" + std::to_string(offset) + ", " +
                              std::to_string(offset + n) + ") are not all loaded");
  if (int rc = set_dev(h)) return rc;
  if (int rc = grow_buf(h->bflag, 16)) return rc;
  unsigned long long *cnt = (unsigned long long *)h->bflag.p;
  HIPX(hipMemsetAsync(cnt, 0, 8, h->stream));
  if (n > 0)
    hipLaunchKernelGGL(refuse_births_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       h->stream, fk, E->k, offset, n, cnt);
  HIPX(hipGetLastError());
  unsigned long long c = 0;
  HIPX(hipMemcpyAsync(&c, cnt, 8, hipMemcpyDeviceToHost, h->stream));
  HIPX(hipStreamSynchronize(h->stream));
  if (n_refused) *n_refused = (int64_t)c;
  return 0;
}

int ig_last_fly_ms(ig_engine *h, double *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_last_fly_ms: null argument");
  if (!h->timed) return fail(IG_E_STATE, "ig_last_fly_ms: no flight launched yet");
  if (int rc = set_dev(h)) return rc;
  HIPX(hipEventSynchronize(h->ev1));
  float ms = 0.0f;
  HIPX(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *out = ms;
  return 0;
}

int ig_fly_batch(ig_engine *h, ig_fields_id f, const ig_flight_params *p, const ig_ions *ions,
                 ig_results *out) {
  if (!h || !p || !ions || !out) return fail(IG_E_ARG, "ig_fly_batch: null argument");
  auto fit = h->fields.find(f);
  if (fit == h->fields.end()) return fail(IG_E_STATE, "ig_fly_batch: unknown field id");
  const int32_t route = fit->second->k.route;
  int32_t cols = 0;
  if (p->rec_rows > 0) {
    if (route == IG_ROUTE_RZ) {
      cols = 7;
      for (int b = 0; b < 13; ++b) cols += (p->rec_mask >> b) & 1u;
    } else if (route == IG_ROUTE_PLANAR) {
      // build_planar.py:1162-1214: the r-z channels without path and the
      // integrals, plus e_x, e_y, e_z
      const uint32_t planar_bits = 0x3FFu | IG_REC_E_X | IG_REC_E_Y | IG_REC_E_Z;
      if (p->rec_mask & ~planar_bits)
        return fail(IG_E_ARG, "ig_fly_batch: record channel not produced by the planar route");
      cols = 7;
      for (int b = 0; b < 16; ++b) cols += (p->rec_mask >> b) & 1u;
    } else {
      cols = p->tp_axis > 0 ? 11 : 10;
    }
  }
  ig_ensemble_id e = 0;
  int rc = ig_ensemble_create(h, ions->n, p->rec_rows, cols, &e);
  if (rc) return rc;
  rc = ig_ensemble_load(h, e, ions);
  if (!rc) {
    ig_flight_params q = *p;
    q.max_steps = 0; // to fate
    rc = ig_fly(h, f, &q, e);
  }
  if (!rc) rc = ig_ensemble_read(h, e, out);
  std::string keep = g_err;
  ig_ensemble_release(h, e);
  if (rc) g_err = keep;
  return rc;
}

int ig_solve3d(ig_engine *h, const ig_solve3d_desc *d, double *phi, ig_solve3d_result *res) {
  if (!h || !d || !phi || !res || !d->fixed || !d->val)
    return fail(IG_E_ARG, "ig_solve3d: null argument");
  for (int a = 0; a < 3; ++a)
    if (d->n[a] < 1 || d->n[a] > 65535)
      return fail(IG_E_ARG, "ig_solve3d: grid sizes must be in [1, 65535]");
  if ((double)d->n[0] * d->n[1] * d->n[2] > 2.0e9)
    return fail(IG_E_ARG, "ig_solve3d: more than 2e9 nodes");
  if (d->method != IG_SOLVE_SOR && d->method != IG_SOLVE_VCYCLE && d->method != IG_SOLVE_CASCADE)
    return fail(IG_E_ARG,
                "ig_solve3d: method must be IG_SOLVE_SOR, IG_SOLVE_VCYCLE or IG_SOLVE_CASCADE");
  if (d->stencil != IG_STENCIL_GHOST_LINEAR && d->stencil != IG_STENCIL_SW)
    return fail(IG_E_ARG, "ig_solve3d: unknown stencil (ghost_linear 0, sw 1)");
  if (!(d->tol > 0.0) || !std::isfinite(d->tol))
    return fail(IG_E_ARG, "ig_solve3d: tol must be positive and finite");
  if (d->method != IG_SOLVE_VCYCLE) {
    if (!(d->omega > 0.0 && d->omega < 2.0))
      return fail(IG_E_ARG, "ig_solve3d: SOR omega must lie in (0, 2)");
    if (d->max_sweeps < 0) return fail(IG_E_ARG, "ig_solve3d: max_sweeps < 0");
    if (d->method == IG_SOLVE_CASCADE) {
      if (d->phi0) return fail(IG_E_ARG, "ig_solve3d: the cascade starts from zeros (phi0 NULL)");
      if (!(d->level_tol_factor > 0.0) || !std::isfinite(d->level_tol_factor))
        return fail(IG_E_ARG, "ig_solve3d: level_tol_factor must be positive and finite");
    }
  } else {
    if (d->max_cycles < 0 || d->nu1 < 0 || d->nu2 < 0)
      return fail(IG_E_ARG, "ig_solve3d: max_cycles, nu1 and nu2 must be >= 0");
    if (d->sym_axes < 0 || d->sym_axes > 7)
      return fail(IG_E_ARG, "ig_solve3d: sym_axes is a 3-bit mask");
    if (std::min(d->n[0], std::min(d->n[1], d->n[2])) < 2)
      return fail(IG_E_ARG, "ig_solve3d: the V-cycle needs every axis of at least 2 nodes");
  }
  if (int rc = set_dev(h)) return rc;
  std::string err;
  if (int rc = ig::solve3d_run(h->stream, d, phi, res, err))
    return fail(rc, "ig_solve3d: " + err);
  return 0;
}

int ig_solve2d(ig_engine *h, const ig_solve2d_desc *d, double *phi, ig_solve3d_result *res) {
  if (!h || !d || !phi || !res || !d->fixed || !d->val)
    return fail(IG_E_ARG, "ig_solve2d: null argument");
  if (d->nz < 2 || d->nz > 65535 || d->nu < 2 || d->nu > 65535)
    return fail(IG_E_ARG, "ig_solve2d: grid sizes must be in [2, 65535]");
  if (d->symmetry != IG_SYM_PLANAR && d->symmetry != IG_SYM_CYLINDRICAL)
    return fail(IG_E_ARG, "ig_solve2d: symmetry must be IG_SYM_PLANAR or IG_SYM_CYLINDRICAL");
  if ((d->neumann & ~15) || (d->ghost_linear & ~15))
    return fail(IG_E_ARG, "ig_solve2d: edge masks are 4-bit (z0 1, z1 2, u0 4, u1 8)");
  if (!(d->h > 0.0) || !std::isfinite(d->h))
    return fail(IG_E_ARG, "ig_solve2d: h must be positive and finite");
  if (!(d->tol > 0.0) || !std::isfinite(d->tol))
    return fail(IG_E_ARG, "ig_solve2d: tol must be positive and finite");
  if (!(d->omega > 0.0 && d->omega < 2.0))
    return fail(IG_E_ARG, "ig_solve2d: SOR omega must lie in (0, 2)");
  if (d->max_sweeps < 0) return fail(IG_E_ARG, "ig_solve2d: max_sweeps < 0");
  // a free node whose missing neighbour has a nonzero coefficient needs
  // its edge open (solver2d.py:104-106); on the cylindrical axis the
  // missing coefficient is zero
  const int64_t nz = d->nz, nu = d->nu;
  const bool cyl = d->symmetry == IG_SYM_CYLINDRICAL;
  const struct {
    int bit;
    const char *name;
  } edges[4] = {{IG_EDGE_Z0, "z0"}, {IG_EDGE_Z1, "z1"}, {IG_EDGE_U0, "u0"}, {IG_EDGE_U1, "u1"}};
  for (const auto &e : edges) {
    if (d->neumann & e.bit) continue;
    if (cyl && e.bit == IG_EDGE_U0) continue;
    const bool zedge = e.bit <= IG_EDGE_Z1;
    const int64_t cnt = zedge ? nu : nz;
    for (int64_t t = 0; t < cnt; ++t) {
      const int64_t i = zedge ? (e.bit == IG_EDGE_Z0 ? 0 : nz - 1) : t;
      const int64_t j = zedge ? t : (e.bit == IG_EDGE_U0 ? 0 : nu - 1);
      if (!d->fixed[i * nu + j])
        return fail(IG_E_ARG, std::string("ig_solve2d: free node with missing neighbour on "
                                          "closed edge ") + e.name);
    }
  }
  if (int rc = set_dev(h)) return rc;
  std::string err;
  if (int rc = ig::solve2d_run(h->stream, d, phi, res, err))
    return fail(rc, "ig_solve2d: " + err);
  return 0;
}

} // extern "C"
