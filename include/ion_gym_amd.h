/*
 * ion_gym_amd.h -- C ABI of the MI355X ion-flight engine.
 *
 * The engine flies ensembles of independent ions through precomputed
 * per-channel E-field grids on one AMD Instinct MI355X (gfx950).  It is
 * the batched replacement of ion_gym's per-ion numba flight kernels:
 *
 *   route RZ  : ion_gym/physics/tracer_rz.py:113-446  (_fly_rec_full)
 *               kick-drift-kick leapfrog, bilinear field on the
 *               mirror-unfolded (axial, radial) half plane, 3-D ion state.
 *   route XYZ : ion_gym/physics/tracer3d.py:280-575   (_fly3d)
 *               and, with SDS gas, tracer3d.py:851-1025 (_fly3d_sds)
 *               classic RK4, trilinear field, 3-D state.
 *
 * The reference has no FFI: its boundary is the Python call
 * fly_fn(i) -> (traj, summary) (ion_gym/physics/build_rz.py:675-776,
 * ion_gym/physics/build_stl3d.py:1181-1396) invoked by
 * ensemble_driver.run / fly_all / run_threaded
 * (ion_gym/physics/ensemble_driver.py:381-529).  The Python package
 * ion_gym_amd binds this header with ctypes (the GIL is released for the
 * duration of every call) and keeps that Python contract unchanged.
 *
 * Conventions
 *  - Every function returns an int status: 0 = ok, < 0 = error class
 *    (IG_E_*).  ig_last_error() returns the calling thread's last message.
 *  - Host buffers passed in are caller owned and copied before return.
 *    Output buffers are caller allocated.  Device memory belongs to the
 *    engine and is freed by the matching *_release call or by
 *    ig_engine_destroy.
 *  - An engine is one device plus one HIP stream.  A handle is used from
 *    one host thread at a time; several engines may share a device.
 *  - Units follow the reference kernels: mm, microseconds, amu, V/m for
 *    r-z channel grids (ion_gym/physics/ionbench.py:92-93), V/mm for 3-D
 *    channel grids (ion_gym/physics/tracer3d.py:295-300).
 */
#ifndef ION_GYM_AMD_H
#define ION_GYM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IG_ABI_VERSION 2

/* status codes */
#define IG_OK 0
#define IG_E_ARG (-1)      /* invalid argument or shape */
#define IG_E_HIP (-2)      /* HIP runtime error */
#define IG_E_NOMEM (-3)    /* device allocation failed */
#define IG_E_STATE (-4)    /* wrong object state (unknown id, busy) */
#define IG_E_KERNEL (-5)   /* a flight refused inside the kernel */

/* routes */
#define IG_ROUTE_RZ 0
#define IG_ROUTE_XYZ 1
#define IG_ROUTE_PLANAR 2 /* ion_gym/physics/build_planar.py:1111-1429 */

/* device storage / arithmetic of the channel grids */
#define IG_PREC_F64 0       /* FP64 storage, FP64 interpolation */
#define IG_PREC_F32_STORE 1 /* FP32 storage, FP64 interpolation: the
                               reference's channel_dtype="float32"
                               (ion_gym/io/sim_spec.py:1339-1340) */
#define IG_PREC_F32 2       /* FP32 storage, FP32 interpolation; state,
                               time and waveform phase stay FP64 */

/* waveform kinds (ion_gym/physics/build_planar.py:1059-1096,
   ion_gym/physics/tracer3d.py:234-276) */
#define IG_WAVE_SIN 0
#define IG_WAVE_COS 1
#define IG_WAVE_SQUARE 2
#define IG_WAVE_TAB_HOLD 3
#define IG_WAVE_TAB_LINEAR 4

/* gas models */
#define IG_GAS_VACUUM 0
#define IG_GAS_HS 1  /* ion_gym/physics/collision3d.py:333-374 */
#define IG_GAS_SDS 2 /* ion_gym/physics/sds.py:134-169, tracer_rz.py:219-272 */

/* fates (ion_gym/physics/ion_envelope.py:100-103) plus the engine's own
   "still flying" marker for windowed flights */
#define IG_FATE_EMPTY (-4)  /* ensemble slot holding no ion */
#define IG_FATE_FLYING (-2)
#define IG_FATE_IMPACT 0
#define IG_FATE_BOX 1
#define IG_FATE_TIMEOUT 2
#define IG_FATE_BOUND 3
#define IG_FATE_TRANSPORTER 4
#define IG_FATE_PLANE 5
#define IG_FATE_DETECT 6

/* ig_results.face of an ion refused at birth (ig_ensemble_refuse_births):
   fate IG_FATE_IMPACT inside metal, IG_FATE_BOX outside the domain */
#define IG_FACE_BIRTH (-2)

/* optional record channels, bit i = the i-th r-z channel of
   ion_gym/physics/build_rz.py:569-571 in that order */
#define IG_REC_SPEED (1u << 0)
#define IG_REC_KE (1u << 1)
#define IG_REC_KE_X (1u << 2)
#define IG_REC_KE_Y (1u << 3)
#define IG_REC_KE_Z (1u << 4)
#define IG_REC_E_FIELD (1u << 5)
#define IG_REC_E_AXIAL (1u << 6)
#define IG_REC_E_RADIAL (1u << 7)
#define IG_REC_RADIUS (1u << 8)
#define IG_REC_N_COL (1u << 9)
#define IG_REC_PATH (1u << 10)
#define IG_REC_E_AXIAL_TINT (1u << 11)
#define IG_REC_KE_TINT (1u << 12)
#define IG_REC_E_X (1u << 13) /* planar route only */
#define IG_REC_E_Y (1u << 14)
#define IG_REC_E_Z (1u << 15)

#define IG_MAX_CHANNELS 16
#define IG_MAX_PLANES 16

typedef struct ig_engine ig_engine;
typedef int64_t ig_fields_id;
typedef int64_t ig_ensemble_id;
typedef int64_t ig_bases_id;

/* Field pack.  Replaces the arrays the reference passes to every kernel
 * call: r-z EzA/EuA/EzK/EuK/EzG/EuG + mirrored ele
 * (ion_gym/physics/build_rz.py:725-738); 3-D EAx..EzK + ele
 * (ion_gym/physics/tracer3d.py:281-290).
 *
 * grids: FP64, C-contiguous, shape (n_sets, n_comp, n0, n1, n2) with
 *   set 0 the static field, sets 1..n_chan the drive channels and, when
 *   has_gate, set n_chan+1 the r-z gate field; n_comp = 2 (axial, radial)
 *   for RZ and 3 (x, y, z) for XYZ; n2 = 1 for RZ.  n1 of an RZ pack is
 *   the mirror-unfolded radial width 2*nr-1.
 * metal: uint8 (n0, n1, n2), nonzero = metal node.  RZ uses the
 *   nearest-node rule (ion_gym/physics/raster2d.py:597-612), XYZ the
 *   all-eight-corners rule (ion_gym/physics/tracer3d.py:174-189).
 * label: optional uint8 (n0, n1, n2) electrode labels (0 = vacuum); the
 *   splat electrode of an impact is label[nearest node of the end point].
 */
typedef struct {
  int32_t route;     /* IG_ROUTE_* */
  int32_t precision; /* IG_PREC_* device storage/arithmetic */
  int64_t n0, n1, n2;
  double h_mm;       /* node pitch */
  double u0;         /* RZ: radial coordinate of column 0, -(nr-1)*h */
  int32_t n_chan;    /* K time-dependent channels, <= IG_MAX_CHANNELS */
  int32_t has_gate;  /* RZ gate set present */
  double tau_gate;   /* RZ gate time (us), < 0 = no gate */
  const double *grids;
  const double *chan_scale; /* optional (n_sets): FP32 storage stores
                               grid/scale and multiplies back; NULL = 1 */
  const uint8_t *metal;
  const uint8_t *label;
  /* waveforms, one entry per drive channel */
  const int32_t *ch_kind;
  const double *ch_om;   /* rad/us */
  const double *ch_ph;   /* rad */
  const double *ch_duty;
  const double *ch_amp;  /* XYZ affine amplitude (r-z bakes it: pass 1) */
  const double *ch_off;  /* XYZ affine offset (r-z: 0) */
  const double *tab_t;   /* shared breakpoints (us) */
  const double *tab_v;
  const int64_t *tab_off; /* (n_chan+1) slice offsets into tab_t/tab_v */
  int64_t n_tab;
  /* IG_FOLD_LINKED: for an IG_PREC_F32 pack, store drive channels whose
     waveforms are linked -- the same kind (sin, cos, or square at duty
     0.5), the same frequency, phases a whole number of half periods
     apart -- as one set.  w(x + n pi) = (-1)^n w(x) for these waveforms,
     so sum_k (a_k w(om t + ph_k) + b_k) E_k = w(om t + ph_L) sum_k
     (-1)^n_k a_k E_k + sum_k b_k E_k: one set per linked group (its grid
     combined in FP64 before the FP32 rounding) plus one constant set when
     an offset b_k is nonzero.  Fewer sets mean fewer gathered bytes and
     interpolation FMAs per step; the field equals the unfolded one to the
     FP32 rounding.  Applied only where it lowers the set count; never to
     packs with table waveforms.  Tunes flown on a folded pack must keep
     each group linked with the pack's amplitude ratios (ig_fly refuses
     others).  A folded pack cannot be composed into (ig_fields_compose);
     ig_fields_layout reports its stored sets.  0 (IG_FOLD_NONE) keeps
     every set as given. */
  int32_t fold_channels;
} ig_fields_desc;

#define IG_FOLD_NONE 0
#define IG_FOLD_LINKED 1

/* Flight parameters shared by every ion of one ig_fly call.  Replaces the
 * scalar/array arguments of _fly_rec_full (ion_gym/physics/tracer_rz.py:
 * 114-123) and _fly3d (ion_gym/physics/tracer3d.py:281-290). */
typedef struct {
  double dt;           /* step in the route's clock: RZ and PLANAR us,
                          XYZ s (the reference's dt_ns*1e-3 and
                          dt_ns*1e-9); XYZ with IG_GAS_SDS us, the clock
                          of _fly3d_sds (tracer3d.py:1117: dt_ns*1e-3) */
  double t_max;        /* flight time limit, same unit, while t < t_max */
  int64_t max_steps;   /* steps this call may advance each ion; <= 0 =
                          until fate.  Ions still flying keep fate
                          IG_FATE_FLYING and resume on the next call. */
  int32_t rec_every;   /* record cadence (steps), >= 1 */
  int32_t rec_rows;    /* rows per ion in the record buffer; 0 = none */
  uint32_t rec_mask;   /* IG_REC_* optional channels (RZ) */
  int32_t gas_model;   /* IG_GAS_* */
  double T_k, P_pa, c_star, c_bar, sig1d, m_gas; /* gas scalars
                          (ion_gym/physics/collision3d.py:207-245) */
  double flow[3];      /* gas flow mm/us */
  const double *sds_stats; /* (5, 1002) jump table for IG_GAS_SDS */
  int32_t bnd_on[6];   /* RZ bounding faces x-,x+,y-,y+,z-,z+ */
  double bnd_val[6];
  int32_t n_planes;    /* station / bounding planes */
  int32_t pl_col[IG_MAX_PLANES];
  double pl_val[IG_MAX_PLANES];
  double pl_sgn[IG_MAX_PLANES];
  double pl_w[IG_MAX_PLANES][4];
  int32_t pl_kind[IG_MAX_PLANES];
  uint64_t rng_key;    /* mixed into every ion's Philox key; lets a
                          serving loop keep requests' streams apart */
  int32_t coll_batch;  /* hard-sphere collisions a 64-lane wave gathers
                          before scattering them together (1 = at once;
                          0 = engine default); results do not depend on
                          it */
  int32_t occupancy;   /* register-budget variant of the flight
                          kernel: at least 1, 2 or 3 waves per SIMD, or 4
                          for the planar FP32-basis vacuum kernel (others
                          take 3); 0 = the engine's measured default per
                          precision, route and gas; results do not depend
                          on it */
  int32_t block;       /* lanes per workgroup: 64, 128, 192 or 256
                          (0 = the engine's choice: 256, or 64 for an
                          ensemble of fewer than eight 64-lane waves per
                          compute unit); results do not depend on it */
  int32_t compact;     /* 1: gather the ensemble's live ions into a dense
                          index list (one wave ballot each) before the
                          window, so dead and empty slots take no lanes;
                          2: the same list in cell order (a counting sort
                          by the grid cell each ion sits in), so the
                          lanes of a wave gather neighbouring field
                          records; 3: the engine's choice per route (cell
                          order for XYZ, slot order otherwise); results
                          do not depend on it */
  /* the periodic transporter of the 3-D RK4 flight (tracer3d.py:433-458,
     fly3d's transporter argument, :722-735): tp_axis 0 = none, else
     1 + the axis (1 x, 2 y, 3 z); an ion crossing the accept plane in tp_direction within a
     step is shifted, with the step's start point, by emit - accept along
     the axis; the tp_max_passes-th crossing ends it with fate
     IG_FATE_TRANSPORTER.  Records then carry an 11th column, the pass
     count (the reference's wrap_passes). */
  int32_t tp_axis;
  double tp_accept_mm, tp_emit_mm, tp_direction;
  int64_t tp_max_passes;
} ig_flight_params;

/* One tune: the parameters one request or ensemble of a batched launch
 * flies with.  An ensemble with a tune table flies each ion with its own
 * tune's waveforms, field-set scales, step, time limit and RNG key
 * instead of the field pack's waveforms and the flight's dt/t_max/rng_key
 * (all else comes from the flight).  This is how a single launch carries
 * many ensembles of different drive tune -- the per-deck re-weighting the
 * reference does by rebuilding its fly function per tune
 * (ion_gym/physics/flight_conditions.py:55-243). */
typedef struct {
  double dt, t_max;   /* route clock, as in ig_flight_params */
  uint64_t rng_key;
  double set_scale[IG_MAX_CHANNELS + 2]; /* multiplies the interpolated
                         field of set s (static, drive channels, gate) */
  int32_t ch_kind[IG_MAX_CHANNELS];      /* sin, cos or square */
  double ch_om[IG_MAX_CHANNELS];
  double ch_ph[IG_MAX_CHANNELS];
  double ch_duty[IG_MAX_CHANNELS];
  double ch_amp[IG_MAX_CHANNELS];
  double ch_off[IG_MAX_CHANNELS];
} ig_tune;

/* Per-ion inputs (host, SoA, n entries each).  Births are the rows of
 * generate_births (ion_gym/physics/sim_build.py:82-84); per-ion scalars
 * are what fly_fn derives (ion_gym/physics/build_rz.py:703-713):
 * acc = q e/(m amu) for RZ, qm = q e/(m amu) for XYZ (same number). */
typedef struct {
  int64_t n;
  const double *x, *y, *z, *vx, *vy, *vz, *tob; /* mm, mm/us, us; all
                          seven NULL: zero, for ig_ensemble_births_at */
  const double *m_amu;
  const double *acc;     /* C/kg */
  const double *sigma_m2;
  const int64_t *seed;   /* per-ion RNG seed (seed_base + i) */
  const double *sds;     /* optional (n, 4): damping, mfp_mm, V_mm_us,
                            log_mr (ion_gym/physics/sds.py:428-447) */
  const int32_t *fate_in; /* optional: a fate preset by the host (birth
                             splat); such ions are not flown */
} ig_ions;

/* Per-ion results (host, caller allocated, SoA). */
typedef struct {
  double *x, *y, *z, *vx, *vy, *vz; /* end state, mm and mm/us */
  double *t_us;          /* lab time of the end state (tob + t) */
  int32_t *fate;         /* IG_FATE_* */
  int32_t *face;         /* RZ bounding face, XYZ hit plane, -1 none */
  int32_t *label;        /* splat electrode label (impact), else 0 */
  int64_t *ncol;         /* hard-sphere collisions */
  int64_t *steps;        /* integrator steps flown */
  int32_t *nrec;         /* record rows written (if records) */
  double *records;       /* optional (n, rec_rows, rec_cols) */
} ig_results;

/* Per-electrode potential bases kept on the device for field composition.
 * Route RZ: the half grid (n0 = nz, n1 = nr, n2 = 1) of electrode
 * potentials solved at base_volts (the reference's cache convention, fa =
 * phi / base_volts, build_rz.py:377-386, 435-534); coord0 = Z_mm (n0) and
 * coord1 = the mirror-unfolded U_mm (2 n1 - 1), the coordinate arrays
 * np.gradient receives (ionbench.py:64-72).  Route XYZ: (nx, ny, nz) unit
 * potentials with uniform spacing h_mm (tracer3d.py:93-117).  metal: the
 * node metal mask (RZ half grid), for the electrode-aware one-sided
 * differences. */
typedef struct {
  int32_t route;       /* IG_ROUTE_* */
  int64_t n0, n1, n2;
  int32_t n_bases;
  const double *phi;   /* (n_bases, n0, n1, n2) */
  const uint8_t *metal;/* (n0, n1, n2), nonzero = metal */
  double base_volts;   /* RZ 1e4, XYZ 1 */
  const double *coord0, *coord1; /* RZ gradient coordinates (see above) */
  double h_mm;         /* XYZ node spacing */
} ig_bases_desc;

/* --- engine ------------------------------------------------------------ */
const char *ig_last_error(void);
int ig_abi_version(void);
int ig_device_count(int *out);
/* The device's marketing name and compute-unit count (the autotuner keys
 * its measurements on them; no counterpart in the reference). */
int ig_device_info(int device, char *name, int32_t name_len, int32_t *cu_count);
int ig_engine_create(int device, ig_engine **out);
int ig_engine_destroy(ig_engine *h);
int ig_sync(ig_engine *h);

/* --- field packs --------------------------------------------------------
 * Upload, convert and interleave a field pack into device memory.
 * Replaces the per-call array arguments listed at ig_fields_desc. */
int ig_fields_upload(ig_engine *h, const ig_fields_desc *d, ig_fields_id *out);
int ig_fields_release(ig_engine *h, ig_fields_id id);
int ig_fields_bytes(ig_engine *h, ig_fields_id id, int64_t *out);

/* --- field composition on the device ------------------------------------
 * Upload per-electrode bases once; then compose any set of a resident
 * field pack as phi = sum over terms t, in order, of w_t * (phi_{b_t} /
 * base_volts) and E = -grad(phi) with np.gradient's coefficients and the
 * electrode-aware one-sided difference into vacuum at metal nodes, V/m on
 * the mirror-unfolded r-z grid or V/mm in 3-D, stored in the pack's
 * precision and per-set scale.  Replaces the host composition and
 * gradient pass of a re-tune: build_rz.py:428-534 with
 * ionbench.build_field_aware (ionbench.py:44-96), tracer3d.py:54-117
 * (build_field_aware_3d, _fa3d_correct). */
int ig_bases_upload(ig_engine *h, const ig_bases_desc *d, ig_bases_id *out);
int ig_bases_release(ig_engine *h, ig_bases_id id);
int ig_fields_compose(ig_engine *h, ig_bases_id b, ig_fields_id f, int32_t set,
                      int32_t n_terms, const int32_t *base_index, const double *weight);
/* Read one component grid of one set of a resident pack back to the host
 * as doubles (stored value x the set's storage scale): (n0, n1, n2).
 * No counterpart in the reference (inspection of device packs). */
int ig_fields_read_set(ig_engine *h, ig_fields_id f, int32_t set, int32_t comp, double *out);
/* The stored layout of a resident pack: its field sets (static, drive
 * channels, gate; after IG_FOLD_LINKED folding), its drive channels, the
 * values per node record, and the drive channels it was given.  No
 * counterpart in the reference (inspection of device packs). */
int ig_fields_layout(ig_engine *h, ig_fields_id f, int32_t *n_sets, int32_t *n_chan,
                     int32_t *stride, int32_t *n_chan_given);

/* --- device-resident ensembles ------------------------------------------
 * Ion state kept in HBM across calls so that long flights, throughput
 * windows and the serving loop never round-trip it through the host.
 * Replaces the per-ion state the reference keeps inside one kernel call
 * (ion_gym/physics/tracer_rz.py:165-206). */
int ig_ensemble_create(ig_engine *h, int64_t capacity, int32_t rec_rows,
                       int32_t rec_cols, ig_ensemble_id *out);
int ig_ensemble_release(ig_engine *h, ig_ensemble_id id);
int ig_ensemble_load(ig_engine *h, ig_ensemble_id id, const ig_ions *ions);
int ig_ensemble_read(ig_engine *h, ig_ensemble_id id, ig_results *out);
/* Load ions into slots [offset, offset + n) with tune index `tune`,
 * leaving the other slots as they are (admission of a request into a
 * running batch).  The kernel-level form of flying ions whose per-ion
 * envelope (ion_gym/physics/build_rz.py:675-738) differs per request. */
int ig_ensemble_load_at(ig_engine *h, ig_ensemble_id id, int64_t offset,
                        const ig_ions *ions, int32_t tune);
/* Read slots [offset, offset + n). */
int ig_ensemble_read_at(ig_engine *h, ig_ensemble_id id, int64_t offset,
                        int64_t n, ig_results *out);
/* Install the ensemble's tune table (n = 0 removes it). */
int ig_ensemble_set_tunes(ig_engine *h, ig_ensemble_id id, int32_t n,
                          const ig_tune *tunes);
/* Ions still flying per tune index, out[0..n) (synchronous). */
int ig_ensemble_live_by_tune(ig_engine *h, ig_ensemble_id id, int32_t n,
                             int64_t *out);
/* Counts ions whose fate is still IG_FATE_FLYING (synchronous). */
int ig_ensemble_live(ig_engine *h, ig_ensemble_id id, int64_t *out);
/* Total integrator steps flown by the ensemble so far (synchronous). */
int ig_ensemble_steps(ig_engine *h, ig_ensemble_id id, int64_t *out);

/* Advance every live ion of the ensemble (asynchronous on the engine's
 * stream).  The kernel-level replacement of one _fly_rec_full /
 * _fly3d call per ion (ion_gym/physics/tracer_rz.py:113,
 * ion_gym/physics/tracer3d.py:280). */
int ig_fly(ig_engine *h, ig_fields_id f, const ig_flight_params *p,
           ig_ensemble_id e);

/* One-shot: load, fly to fate, read.  The batched form of
 * fly_fn(i) over a list of ions (ion_gym/physics/build_rz.py:675). */
int ig_fly_batch(ig_engine *h, ig_fields_id f, const ig_flight_params *p,
                 const ig_ions *ions, ig_results *out);

/* n_windows consecutive ig_fly windows captured once into a HIP graph and
 * replayed: the first call with given kernel arguments (field pack, flight
 * parameters, ensemble state pointers, size, block, n_windows) captures,
 * later calls replay the cached graph; a changed argument captures anew.
 * Equivalent to n_windows ig_fly calls with no host work between them
 * (the window loop of a long flight), with one launch instead of
 * n_windows.  The graph holds only the flight kernels (and, with
 * compact, the live-list kernels before each window).  Replaces the
 * host loop over windows (ion_gym/physics/ensemble_driver.py:381-529). */
int ig_fly_graph(ig_engine *h, ig_fields_id f, const ig_flight_params *p,
                 ig_ensemble_id e, int32_t n_windows);

/* Kernel timing of the last ig_fly on the engine's stream (HIP events
 * recorded around the launches), in milliseconds. */
int ig_last_fly_ms(ig_engine *h, double *out);

/* ig_fly without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate while they launch (code under HIP
 * graph capture, serving loops and harnesses that forbid hipMalloc).
 * Results equal ig_fly's bit for bit.  The live-ion list, its cell order
 * and the SDS jump table of the window live in one fixed region of device
 * memory per GPU (a static device array of about 14 MB, held until the
 * process ends), set up when an engine is created on that device; the
 * call allocates nothing (the ensemble and the field pack are the
 * caller's, created beforehand).  At most IG_NOALLOC_MAX_IONS ions per
 * ensemble (IG_E_ARG beyond).  Every call on a device shares its region:
 * calls that overlap in time on one device race.  Two inputs are copied
 * to the device (host work, then a synchronizing copy, no allocation) when
 * they change: the SDS jump table, and an ensemble's tune table re-laid for
 * an IG_FOLD_LINKED pack it has not flown on last; a caller capturing the
 * call into a graph flies once before capturing. */
#define IG_NOALLOC_MAX_IONS (1 << 20)
int ig_fly_noalloc(ig_engine *h, ig_fields_id f, const ig_flight_params *p,
                   ig_ensemble_id e);

/* ---- births on the device ------------------------------------------- */

#define IG_BIRTH_POINT 0
#define IG_BIRTH_DISC 1
#define IG_BIRTH_LINE 2
#define IG_BIRTH_GRID 3
#define IG_BIRTH_BOX 4
#define IG_BIRTH_GAUSSIAN 5

#define IG_E_BIRTH_TRIES (-6) /* a truncated-Gaussian birth drew no value
                                 inside its truncation in gauss_tries */

/* One source of the reference's generate_births
 * (ion_gym/physics/sim_build.py:82-277), for every distribution but
 * "file" (rows read from a CSV on the host).  The n_per * n_mz ions draw,
 * in index order, from ONE numpy PCG64 stream (np.random.default_rng(seed))
 * exactly as the reference's loop does: the position draws of the
 * distribution, then the kinetic energy, the thermal and the per-axis
 * velocity spreads, then the birth time.  Host-derived quantities are
 * passed as the reference computes them in Python. */
typedef struct {
  int32_t distribution;   /* IG_BIRTH_* */
  int32_t axis;           /* disc normal / line direction: 0 x, 1 y, 2 z */
  int64_t n_per;          /* ions per mass (source.n_ions) */
  int32_t n_mz;           /* masses in mz_list (>= 1) */
  const double *mz_list;  /* n_mz masses, Da (host) */
  double p0[3];           /* x0_mm, y0_mm, z0_mm */
  double r_mm, len_mm;
  double box_mm[3];
  double gauss_sigma[3];  /* gaussian_layout(): sigma per axis, mm (0 = none) */
  double gauss_trunc[3];  /* truncation half-widths, mm */
  int32_t gauss_tries;    /* GAUSSIAN_BIRTH_TRIES (sim_build.py:78) */
  int32_t has_ke;         /* ke_hi > 0 or ke_lo > 0: an energy is drawn */
  double ke_lo, ke_hi;    /* eV */
  double dir_unit[3];     /* direction / |direction| (numpy), or zeros */
  int32_t has_dir;        /* |direction| > 0: the directed term is added */
  double temperature_k;   /* > 0: thermal spread sqrt(kT/m)/1000 */
  int32_t has_dv;         /* any dv_fwhm_ms entry nonzero */
  double dv_sigma[3];     /* dv_fwhm_ms * 1e-3 / FWHM_PER_SIGMA, mm/us */
  int32_t has_drift;      /* any v_drift_mm_us entry nonzero */
  double v_drift[3];      /* mm/us */
  double tob_span_us;
  uint64_t pcg_state[2];  /* numpy PCG64 state after seeding: (high, low) */
  uint64_t pcg_inc[2];    /* its increment: (high, low) */
  const uint64_t *zig_k;  /* numpy's standard-normal ziggurat: 256 fast-  */
  const double *zig_w;    /* accept thresholds, layer widths and layer    */
  const double *zig_f;    /* densities (host) */
} ig_births_desc;

/* The (N, 7) births [x, y, z, vx, vy, vz, tob] of a source into a host
 * buffer (row-major, caller allocated), N = n_per * n_mz: the replacement
 * of generate_births (ion_gym/physics/sim_build.py:82-277).  Returns
 * IG_E_BIRTH_TRIES where the reference raises for a truncated Gaussian
 * that no draw satisfies (sim_build.py:190-205). */
int ig_births(ig_engine *h, const ig_births_desc *d, double *out);

/* The same births written on the device into the birth columns (x..vz,
 * tob) of ensemble slots [offset, offset + N), which must already be
 * loaded (ig_ensemble_load_at with NULL birth pointers leaves them zero
 * for this call): a request's ions never pass through host memory.
 * Replaces generate_births ahead of fly_fn for a served request
 * (ion_gym/physics/sim_build.py:722 build_run). */
int ig_ensemble_births_at(ig_engine *h, ig_ensemble_id id, int64_t offset,
                          const ig_births_desc *d);

/* The per-ion birth refusals of the fly functions, on the device, for the
 * not yet flown ions of slots [offset, offset + n): a birth outside the
 * solved domain (ion_gym/physics/birth_domain.py:27-38, nodes 0..n-1
 * with a 1e-9 node tolerance) ends with fate IG_FATE_BOX, one inside
 * metal with fate IG_FATE_IMPACT and its electrode label -- by the
 * nearest-node rule on the half grid for RZ and PLANAR
 * (ion_gym/physics/raster2d.py:624-646, r = hypot(y, z) for RZ,
 * build_rz.py:679-690) and the all-eight-corners rule for XYZ
 * (ion_gym/physics/tracer3d.py:192-230).  Refused ions keep their birth
 * state, get face IG_FACE_BIRTH and are never flown, as the reference's
 * birth_splat (ion_gym/physics/ion_envelope.py:184-223).  *n_refused
 * (optional) receives how many were refused. */
int ig_ensemble_refuse_births(ig_engine *h, ig_ensemble_id id,
                              ig_fields_id f, int64_t offset, int64_t n,
                              int64_t *n_refused);

/* ---- 3-D Laplace basis solve (upstream of the flight: the bases) ---- */

#define IG_SOLVE_SOR 0    /* red-black SOR: ion_gym/physics/solver3d.py:250-293 */
#define IG_SOLVE_VCYCLE 1 /* V-cycle multigrid: ion_gym/physics/multigrid3d.py:602-731 */
#define IG_SOLVE_CASCADE 2 /* cascadic multigrid: ion_gym/physics/multigrid3d.py:175-230 */
#define IG_STENCIL_GHOST_LINEAR 0 /* solver3d.py:172-181 */
#define IG_STENCIL_SW 1           /* Shortley-Weller, solver3d.py:182-190 */

typedef struct {
  int64_t n[3];          /* nx, ny, nz; arrays are C order, z fastest */
  const uint8_t *fixed;  /* (nx, ny, nz) Dirichlet mask (host) */
  const double *val;     /* potentials at the fixed nodes (host) */
  const float *theta;    /* (6, nx, ny, nz) leg fractions in the order
                            -x, +x, -y, +y, -z, +z, or NULL for all 1 */
  const double *phi0;    /* initial guess or NULL for zeros (NULL for
                            the cascade) */
  int32_t method;        /* IG_SOLVE_* */
  int32_t stencil;       /* IG_STENCIL_* */
  double tol;            /* volts */
  double omega;          /* SOR relaxation factor (solver3d.optimal_omega);
                            the cascade's fine grid */
  int64_t max_sweeps;    /* SOR and every cascade grid: red-black sweep
                            pairs allowed */
  int32_t max_cycles;    /* V-cycle: cycles allowed (reference 60) */
  int32_t nu1, nu2;      /* V-cycle: pre/post smoothing pairs (3, 3) */
  int32_t sym_axes;      /* V-cycle: bit a projects each cycle's iterate
                            onto the mirror-symmetric subspace of axis a;
                            the caller has verified the symmetry */
  int32_t min_coarse;    /* cascade: stop coarsening at an axis of at most
                            this many nodes (reference 9) */
  int32_t max_levels;    /* cascade: grids in the hierarchy, the fine one
                            included (8) */
  double level_tol_factor; /* cascade: intermediate grids converge to
                              tol times this (30) */
} ig_solve3d_desc;

typedef struct {
  int64_t sweeps;  /* red-black pairs on the fine grid */
  double delta;    /* SOR: the last pair's max |update|; V-cycle: the last
                      cycle's max |change| */
  int32_t status;  /* 0 converged (delta < tol), 1 the sweep or cycle
                      limit was reached, 2 the V-cycle stagnated */
  int32_t level;   /* cascade: the last grid solved (0 the fine one, l
                      the l-th coarsening); with status 1, the grid whose
                      SOR ran out of sweeps, and sweeps and delta are its */
} ig_solve3d_result;

/* Solve Laplace for one basis and write the final iterate to phi (host,
 * n nodes).  SOR is solve3d: pairs of colour sweeps (node colour
 * (i + j + k) & 1, even reflection at every outer face) until a pair's
 * max |update| < tol.  VCYCLE is solve3d_vmg: per cycle nu1 pairs of
 * Gauss-Seidel sweeps of the same stencil, a plain 7-point error equation
 * on a hierarchy of 2x-coarsened grids (fixed where any fine node of the
 * 2x2x2 block is) solved to 5 % of its right-hand side
 * (multigrid3d.py:553-599), its prolongation added, nu2 pairs, the
 * symmetric projection, and convergence when the cycle moved no node by
 * tol.  CASCADE is solve3d_mg: (fixed, val) coarsened by injection with
 * the block's largest-|val| fixed node rescuing a free injection point,
 * while every axis exceeds min_coarse and halves to at least 3 nodes;
 * the coarsest grid solved by SOR from zeros to tol, each finer one from
 * the trilinear prolongation of the coarser solution (to tol times
 * level_tol_factor, unit fractions, the grid's optimal omega), the fine
 * grid last with its fractions, omega and tol.  Replaces solve3d /
 * solve3d_vmg / solve3d_mg under solve_bases and solve_bases_mg
 * (solver3d.py:296-320, multigrid3d.py:233-302); the caller raises the
 * reference's SolveNotConverged on status != 0. */
int ig_solve3d(ig_engine *h, const ig_solve3d_desc *d, double *phi,
               ig_solve3d_result *res);

/* ---- 2-D Laplace solve (r-z and planar bases) ---- */

#define IG_SYM_PLANAR 0      /* phi_zz + phi_uu = 0 */
#define IG_SYM_CYLINDRICAL 1 /* phi_zz + phi_rr + phi_r / r = 0, u = r >= 0 */
#define IG_EDGE_Z0 1
#define IG_EDGE_Z1 2
#define IG_EDGE_U0 4
#define IG_EDGE_U1 8

typedef struct {
  int64_t nz, nu;         /* grid (nz, nu), u fastest; both at least 2 */
  const uint8_t *fixed;   /* Dirichlet mask (host) */
  const double *val;      /* potentials at the fixed nodes (host) */
  const double *phi0;     /* initial guess or NULL for zeros */
  double h;               /* node spacing, both axes */
  int32_t symmetry;       /* IG_SYM_* */
  int32_t neumann;        /* open edges, IG_EDGE_* bits (u0 ignored when
                             cylindrical: the axis needs no ghost) */
  int32_t ghost_linear;   /* open edges whose missing neighbour folds onto
                             the centre node ('ghost_linear'); the others
                             fold onto the opposite neighbour ('mirror') */
  double tol;             /* stop when a red-black pair moves no node by
                             tol volts */
  double omega;           /* SOR relaxation factor in (0, 2) */
  int64_t max_sweeps;     /* red-black pairs allowed */
} ig_solve2d_desc;

/* Solve the 2-D system of solver2d.assemble_laplace
 * (ion_gym/physics/solver2d.py:31-124): planar interior centre -4 with
 * unit neighbours, cylindrical interior radial neighbours 1 +- h/(2r),
 * the axis centre -6 with the first off-axis node 4, open edges folded
 * per edge_ghost, fixed nodes Dirichlet.  The reference factors it by a
 * sparse LU (laplace_factor / solve_laplace, solver2d.py:127-183); here
 * red-black SOR iterates it on the device to tol, the result written to
 * phi (host, nz * nu).  A free node on a closed edge is refused, as the
 * reference's assembly refuses it.  Convergence is measured every 8th
 * pair, so res->sweeps is a multiple of 8 unless the limit ended the run.
 * res->status 0 converged, 1 the sweep limit was reached. */
int ig_solve2d(ig_engine *h, const ig_solve2d_desc *d, double *phi,
               ig_solve3d_result *res);

#ifdef __cplusplus
}
#endif
#endif /* ION_GYM_AMD_H */
