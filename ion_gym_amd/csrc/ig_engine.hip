// C ABI of the MI355X ion-flight engine (include/ion_gym_amd.h): engine,
// field packs, device-resident ensembles and the flight launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ig_common.h"

using namespace ig;

namespace {

thread_local std::string g_err;

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
  ~DevBuf() {
    if (p) (void)hipFree(p);
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

struct Fields {
  FieldsK k{};
  DevBuf grid, metal, label, tab_t, tab_v;
  int64_t bytes = 0;
};

struct Ensemble {
  int64_t cap = 0, n = 0;
  int32_t route = IG_ROUTE_RZ; // clock of k.t: r-z us, 3-D s
  int32_t rec_rows = 0, rec_cols = 0;
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
  int block = 256;
};

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
  h->fields.clear();
  h->ens.clear();
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

int ig_fields_upload(ig_engine *h, const ig_fields_desc *d, ig_fields_id *out) {
  if (!h || !d || !out) return fail(IG_E_ARG, "ig_fields_upload: null argument");
  if (d->route != IG_ROUTE_RZ && d->route != IG_ROUTE_XYZ)
    return fail(IG_E_ARG, "ig_fields_upload: unknown route " + std::to_string(d->route));
  if (d->precision < IG_PREC_F64 || d->precision > IG_PREC_F32)
    return fail(IG_E_ARG, "ig_fields_upload: unknown precision");
  const bool rz = d->route == IG_ROUTE_RZ;
  if (d->n0 < 2 || d->n1 < 2 || (rz ? d->n2 != 1 : d->n2 < 2))
    return fail(IG_E_ARG, "ig_fields_upload: grid shape must be at least 2 nodes per axis "
                          "(n2 == 1 for the r-z route)");
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
  k.h = d->h_mm;
  k.inv_h = 1.0 / d->h_mm;
  k.u0 = d->u0;
  k.K = d->n_chan;
  k.has_gate = rz && d->has_gate ? 1 : 0;
  k.tau_gate = k.has_gate ? d->tau_gate : -1.0;
  k.n_sets = 1 + k.K + k.has_gate;
  k.n_comp = rz ? 2 : 3;
  const int nvals = k.n_sets * k.n_comp;
  k.stride = record_stride(k.prec, nvals);
  for (int s = 0; s < k.n_sets; ++s) {
    double sc = d->chan_scale ? d->chan_scale[s] : 1.0;
    if (!(sc > 0.0) || !std::isfinite(sc))
      return fail(IG_E_ARG, "ig_fields_upload: chan_scale must be finite and > 0");
    k.scale[s] = k.prec == IG_PREC_F64 ? 1.0 : sc;
  }
  int64_t tab_total = 0;
  for (int c = 0; c < k.K; ++c) {
    int kind = d->ch_kind[c];
    if (kind < IG_WAVE_SIN || kind > IG_WAVE_TAB_LINEAR)
      return fail(IG_E_ARG, "ig_fields_upload: unknown waveform kind " + std::to_string(kind));
    k.kind[c] = kind;
    k.om[c] = d->ch_om[c];
    k.ph[c] = d->ch_ph[c];
    k.duty[c] = d->ch_duty[c];
    k.amp[c] = d->ch_amp ? d->ch_amp[c] : 1.0;
    k.off[c] = d->ch_off ? d->ch_off[c] : 0.0;
  }
  for (int c = 0; c <= k.K; ++c) k.tab_off[c] = d->tab_off ? d->tab_off[c] : 0;
  if (d->tab_off) {
    tab_total = d->n_tab;
    for (int c = 0; c < k.K; ++c)
      if (k.tab_off[c] < 0 || k.tab_off[c + 1] < k.tab_off[c] || k.tab_off[c + 1] > tab_total)
        return fail(IG_E_ARG, "ig_fields_upload: tab_off slices out of range");
  }
  for (int c = 0; c < k.K; ++c)
    if ((k.kind[c] == IG_WAVE_TAB_HOLD || k.kind[c] == IG_WAVE_TAB_LINEAR) &&
        k.tab_off[c + 1] > k.tab_off[c] && (!d->tab_t || !d->tab_v))
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
  for (int s = 0; s < k.n_sets; ++s) {
    const double inv = 1.0 / (d->chan_scale ? d->chan_scale[s] : 1.0);
    for (int c = 0; c < k.n_comp; ++c) {
      const double *src = d->grids + ((int64_t)s * k.n_comp + c) * nodes;
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
  HIPX(hipStreamSynchronize(h->stream));
  k.grid = F->grid.p;
  k.metal = (const uint8_t *)F->metal.p;
  k.label = (const uint8_t *)F->label.p;
  k.tab_t = (const double *)F->tab_t.p;
  k.tab_v = (const double *)F->tab_v.p;
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
  int32_t **i32[] = {&k.has_gauss, &k.fate, &k.face, &k.label, &k.nrec};
  for (int32_t **a : i32)
    if ((rc = E->alloc(*a, n))) return rc;
  k.rec = nullptr;
  if (rec_rows > 0 && (rc = E->alloc(k.rec, n * (size_t)rec_rows * rec_cols))) return rc;
  k.idx = nullptr;
  k.n_idx = 0;
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
  h->ens.erase(it);
  return 0;
}

int ig_ensemble_load(ig_engine *h, ig_ensemble_id id, const ig_ions *io) {
  if (!h || !io) return fail(IG_E_ARG, "ig_ensemble_load: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_load: unknown ensemble id");
  if (io->n < 0 || io->n > E->cap)
    return fail(IG_E_ARG, "ig_ensemble_load: n = " + std::to_string(io->n) +
                              " exceeds the ensemble capacity " + std::to_string(E->cap));
  if (io->n > 0 && (!io->x || !io->y || !io->z || !io->vx || !io->vy || !io->vz || !io->tob ||
                    !io->m_amu || !io->acc || !io->seed))
    return fail(IG_E_ARG, "ig_ensemble_load: births, m_amu, acc and seed are required");
  if (int rc = set_dev(h)) return rc;
  const int64_t n = io->n;
  EnsK &k = E->k;
  E->n = n;
  k.n = n;
  if (n == 0) return 0;
  const size_t b8 = (size_t)n * 8, b4 = (size_t)n * 4;
  hipStream_t st = h->stream;
  HIPX(hipMemcpyAsync(k.x, io->x, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.y, io->y, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.z, io->z, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.vx, io->vx, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.vy, io->vy, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.vz, io->vz, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.tob, io->tob, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.m_amu, io->m_amu, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.acc, io->acc, b8, hipMemcpyHostToDevice, st));
  HIPX(hipMemcpyAsync(k.seed, io->seed, b8, hipMemcpyHostToDevice, st));
  if (io->sigma_m2)
    HIPX(hipMemcpyAsync(k.sigma, io->sigma_m2, b8, hipMemcpyHostToDevice, st));
  else
    HIPX(hipMemsetAsync(k.sigma, 0, b8, st));
  if (io->sds) {
    // (n, 4) rows -> four columns
    std::vector<double> col((size_t)n);
    double *dst[4] = {k.sds_damp, k.sds_mfp, k.sds_V, k.sds_logmr};
    for (int c = 0; c < 4; ++c) {
      for (int64_t i = 0; i < n; ++i) col[i] = io->sds[4 * i + c];
      HIPX(hipMemcpyAsync(dst[c], col.data(), b8, hipMemcpyHostToDevice, st));
      HIPX(hipStreamSynchronize(st));
    }
  } else {
    HIPX(hipMemsetAsync(k.sds_damp, 0, b8, st));
    HIPX(hipMemsetAsync(k.sds_mfp, 0, b8, st));
    HIPX(hipMemsetAsync(k.sds_V, 0, b8, st));
    HIPX(hipMemsetAsync(k.sds_logmr, 0, b8, st));
  }
  HIPX(hipMemsetAsync(k.t, 0, b8, st));
  HIPX(hipMemsetAsync(k.rng_n, 0, b8, st));
  HIPX(hipMemsetAsync(k.gauss, 0, b8, st));
  HIPX(hipMemsetAsync(k.has_gauss, 0, b4, st));
  HIPX(hipMemsetAsync(k.ncol, 0, b8, st));
  HIPX(hipMemsetAsync(k.steps, 0, b8, st));
  HIPX(hipMemsetAsync(k.path, 0, b8, st));
  HIPX(hipMemsetAsync(k.eat, 0, b8, st));
  HIPX(hipMemsetAsync(k.ket, 0, b8, st));
  HIPX(hipMemsetAsync(k.e_last, 0, 3 * b8, st));
  HIPX(hipMemsetAsync(k.face, 0xff, b4, st)); // -1
  HIPX(hipMemsetAsync(k.label, 0, b4, st));
  HIPX(hipMemsetAsync(k.nrec, 0, b4, st));
  if (io->fate_in) {
    std::vector<int32_t> fz((size_t)n);
    for (int64_t i = 0; i < n; ++i) fz[i] = io->fate_in[i] == IG_FATE_FLYING ? IG_FATE_FLYING
                                                                              : io->fate_in[i];
    HIPX(hipMemcpyAsync(k.fate, fz.data(), b4, hipMemcpyHostToDevice, st));
    HIPX(hipStreamSynchronize(st));
  } else {
    std::vector<int32_t> fz((size_t)n, IG_FATE_FLYING);
    HIPX(hipMemcpyAsync(k.fate, fz.data(), b4, hipMemcpyHostToDevice, st));
    HIPX(hipStreamSynchronize(st));
  }
  HIPX(hipStreamSynchronize(st));
  return 0;
}

int ig_ensemble_read(ig_engine *h, ig_ensemble_id id, ig_results *o) {
  if (!h || !o) return fail(IG_E_ARG, "ig_ensemble_read: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_read: unknown ensemble id");
  if (int rc = set_dev(h)) return rc;
  const EnsK &k = E->k;
  const int64_t n = E->n;
  if (n == 0) return 0;
  const size_t b8 = (size_t)n * 8, b4 = (size_t)n * 4;
  hipStream_t st = h->stream;
  HIPX(hipStreamSynchronize(st));
  struct {
    void *dst;
    const void *src;
    size_t bytes;
  } cp[] = {{o->x, k.x, b8},         {o->y, k.y, b8},         {o->z, k.z, b8},
            {o->vx, k.vx, b8},       {o->vy, k.vy, b8},       {o->vz, k.vz, b8},
            {o->fate, k.fate, b4},   {o->face, k.face, b4},   {o->label, k.label, b4},
            {o->ncol, k.ncol, b8},   {o->steps, k.steps, b8}, {o->nrec, k.nrec, b4}};
  for (auto &c : cp)
    if (c.dst) HIPX(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, st));
  if (E->route == IG_ROUTE_XYZ && (o->vx || o->vy || o->vz)) {
    // the 3-D kernel carries m/s from its first entry on (nrec >= 1);
    // report mm/us as the reference does (tracer3d.py:779: v * 1e-3)
    std::vector<int32_t> nr((size_t)n);
    HIPX(hipMemcpyAsync(nr.data(), k.nrec, b4, hipMemcpyDeviceToHost, st));
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
    HIPX(hipMemcpyAsync(t.data(), k.t, b8, hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(tob.data(), k.tob, b8, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    if (E->route == IG_ROUTE_XYZ)
      for (int64_t i = 0; i < n; ++i) o->t_us[i] = tob[i] + t[i] * 1e6;
    else
      for (int64_t i = 0; i < n; ++i) o->t_us[i] = tob[i] + t[i];
  }
  if (o->records && E->rec_rows > 0 && k.rec)
    HIPX(hipMemcpyAsync(o->records, k.rec, (size_t)n * E->rec_rows * E->rec_cols * 8,
                        hipMemcpyDeviceToHost, st));
  HIPX(hipStreamSynchronize(st));
  return 0;
}

int ig_ensemble_live(ig_engine *h, ig_ensemble_id id, int64_t *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_ensemble_live: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_live: unknown ensemble id");
  if (int rc = set_dev(h)) return rc;
  std::vector<int32_t> f((size_t)E->n);
  if (E->n) {
    HIPX(hipMemcpyAsync(f.data(), E->k.fate, (size_t)E->n * 4, hipMemcpyDeviceToHost, h->stream));
    HIPX(hipStreamSynchronize(h->stream));
  }
  int64_t live = 0;
  for (int32_t v : f) live += v == IG_FATE_FLYING;
  *out = live;
  return 0;
}

int ig_ensemble_steps(ig_engine *h, ig_ensemble_id id, int64_t *out) {
  if (!h || !out) return fail(IG_E_ARG, "ig_ensemble_steps: null argument");
  Ensemble *E = find_ens(h, id);
  if (!E) return fail(IG_E_STATE, "ig_ensemble_steps: unknown ensemble id");
  if (int rc = set_dev(h)) return rc;
  std::vector<int64_t> s((size_t)E->n);
  if (E->n) {
    HIPX(hipMemcpyAsync(s.data(), E->k.steps, (size_t)E->n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPX(hipStreamSynchronize(h->stream));
  }
  int64_t tot = 0;
  for (int64_t v : s) tot += v;
  *out = tot;
  return 0;
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
  return 0;
}

int ig_fly(ig_engine *h, ig_fields_id fid, const ig_flight_params *p, ig_ensemble_id eid) {
  if (!h || !p) return fail(IG_E_ARG, "ig_fly: null argument");
  auto fit = h->fields.find(fid);
  if (fit == h->fields.end()) return fail(IG_E_STATE, "ig_fly: unknown field id");
  Ensemble *E = find_ens(h, eid);
  if (!E) return fail(IG_E_STATE, "ig_fly: unknown ensemble id");
  FlightK k;
  if (int rc = build_flight(p, k)) return rc;
  if (p->rec_rows > 0 && (p->rec_rows != E->rec_rows || E->k.rec == nullptr))
    return fail(IG_E_ARG, "ig_fly: rec_rows differs from the ensemble's record buffer");
  k.rec_cols = E->rec_cols;
  if (int rc = set_dev(h)) return rc;
  if (k.gas == IG_GAS_SDS) {
    // the (5, 1002) jump table stays on the device until it changes
    const size_t nt = 5 * 1002;
    if (h->sds_host.size() != nt ||
        std::memcmp(h->sds_host.data(), p->sds_stats, nt * sizeof(double)) != 0) {
      HIPX(hipStreamSynchronize(h->stream));
      h->sds_host.assign(p->sds_stats, p->sds_stats + nt);
      if (!h->sds.p) {
        if (int rc = dev_alloc(h->sds, nt * sizeof(double))) return rc;
      }
      HIPX(hipMemcpyAsync(h->sds.p, h->sds_host.data(), nt * sizeof(double),
                          hipMemcpyHostToDevice, h->stream));
      HIPX(hipStreamSynchronize(h->stream));
    }
    k.sds_stats = (const double *)h->sds.p;
  }
  const FieldsK &f = fit->second->k;
  E->route = f.route;
  HIPX(hipEventRecord(h->ev0, h->stream));
  int rc = launch_route(f, k, E->k, E->n, h->block, h->stream);
  HIPX(hipEventRecord(h->ev1, h->stream));
  h->timed = true;
  if (rc == -2) return fail(IG_E_ARG, "ig_fly: no kernel instantiated for this field pack");
  if (rc) {
    hipError_t e = hipGetLastError();
    return fail(IG_E_HIP, std::string("ig_fly: launch failed: ") + hipGetErrorString(e));
  }
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
  const bool rz = fit->second->k.route == IG_ROUTE_RZ;
  int32_t cols = 0;
  if (p->rec_rows > 0) {
    if (rz) {
      cols = 7;
      for (int b = 0; b < 13; ++b) cols += (p->rec_mask >> b) & 1u;
    } else {
      cols = 10;
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

} // extern "C"
