/* C ABI of the MI355X-native MIND simulation step (libmind).
 *
 * The library replaces the reference's device side and its CPU event loop:
 * build! packing, initialize!, run!/step!, spike detection, event routing and
 * delivery, the cable solve and the mechanism phases. A host layer (a
 * stdlib-only Julia module calling these with ccall, or Python ctypes) keeps
 * the reference's model-building API and flattens a model into
 * mind_model_desc. All arrays are plain caller-owned host pointers; the
 * library copies what it needs and owns every device allocation.
 *
 * Every function returns 0 on success or a MIND_E* code; mind_last_error()
 * gives the message of the last failure on the calling thread.
 * Calls on one mind_state must be serialised by the caller; distinct states
 * are independent.
 */
#ifndef MIND_H
#define MIND_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#endif
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. 3, 5, 6 and 8 keep the reference's device error-word meaning
 * (src/runtime/utils/step.jl:183-189). */
enum {
  MIND_OK = 0,
  MIND_EARG = 1,      /* invalid argument (reference: ArgumentError) */
  MIND_EHIP = 2,      /* HIP runtime failure or missing device */
  MIND_EOOM = 3,      /* device allocation failed */
  MIND_EEVENT = 5,    /* invalid self-event delay or target */
  MIND_ESINGULAR = 6, /* singular or non-finite cable system */
  MIND_ENONFINITE = 8,/* non-finite numerical state */
  MIND_EQUEUE = 9,    /* event ring or spike buffer capacity exceeded */
  MIND_EUNSUPPORTED = 10 /* model feature outside the device path */
};

/* Builtin mechanisms (reference mod/{hh,pas,ExpSyn,Exp2Syn,IClamp,NetStim,AlphaSynapse,
 * PatternStim,SEClamp,VClamp}.mod). */
enum {
  MIND_MECH_HH = 0,
  MIND_MECH_PAS = 1,
  MIND_MECH_EXPSYN = 2,
  MIND_MECH_EXP2SYN = 3,
  MIND_MECH_ICLAMP = 4,
  MIND_MECH_NETSTIM = 5,
  MIND_MECH_ALPHASYNAPSE = 6,
  MIND_MECH_PATTERNSTIM = 7, /* artificial cell that emits given spike times (no fields) */
  MIND_MECH_SECLAMP = 8,     /* single-electrode three-level voltage clamp */
  MIND_MECH_VCLAMP = 9,      /* two-electrode three-level voltage clamp (amplifier model) */
  MIND_MECH_COUNT = 10
};

/* Mechanism slots: the builtins above, then mechanisms compiled at run time by
 * mind_compile_mod (ids MIND_MECH_COUNT .. MIND_MECH_SLOTS - 1). */
enum { MIND_MECH_SLOTS = 64 };

/* Field columns of each builtin mechanism (parameters, states, assigned). */
enum { HH_GNABAR, HH_GKBAR, HH_GL, HH_EL, HH_M, HH_H, HH_N, HH_ENA, HH_EK, HH_GNA, HH_GK,
       HH_INA, HH_IK, HH_IL, HH_MINF, HH_HINF, HH_NINF, HH_MTAU, HH_HTAU, HH_NTAU, HH_NFIELDS };
enum { PAS_G, PAS_E, PAS_I, PAS_NFIELDS };
enum { EXPSYN_TAU, EXPSYN_E, EXPSYN_G, EXPSYN_I, EXPSYN_NFIELDS };
enum { EXP2SYN_TAU1, EXP2SYN_TAU2, EXP2SYN_E, EXP2SYN_A, EXP2SYN_B, EXP2SYN_I, EXP2SYN_G,
       EXP2SYN_FACTOR, EXP2SYN_NFIELDS };
enum { ICLAMP_DEL, ICLAMP_DUR, ICLAMP_AMP, ICLAMP_I, ICLAMP_NFIELDS };
enum { NETSTIM_INTERVAL, NETSTIM_NUMBER, NETSTIM_START, NETSTIM_NOISE, NETSTIM_EVENT,
       NETSTIM_ON, NETSTIM_ISPIKE, NETSTIM_NFIELDS };
enum { ALPHASYNAPSE_ONSET, ALPHASYNAPSE_TAU, ALPHASYNAPSE_GMAX, ALPHASYNAPSE_E, ALPHASYNAPSE_I,
       ALPHASYNAPSE_G, ALPHASYNAPSE_NFIELDS };
enum { PATTERNSTIM_NFIELDS = 0 };
enum { SECLAMP_RS, SECLAMP_DUR1, SECLAMP_AMP1, SECLAMP_DUR2, SECLAMP_AMP2, SECLAMP_DUR3, SECLAMP_AMP3,
       SECLAMP_I, SECLAMP_VC, SECLAMP_TC2, SECLAMP_TC3, SECLAMP_ON, SECLAMP_NFIELDS };
/* dur[3] and amp[3] are arrays: element k of dur is column VCLAMP_DUR + k. */
enum { VCLAMP_DUR = 0, VCLAMP_AMP = 3, VCLAMP_GAIN = 6, VCLAMP_RSTIM, VCLAMP_TAU1, VCLAMP_TAU2, VCLAMP_E0,
       VCLAMP_VO0, VCLAMP_VI0, VCLAMP_FAC, VCLAMP_E, VCLAMP_VO, VCLAMP_VI, VCLAMP_I, VCLAMP_STIM, VCLAMP_TC,
       VCLAMP_NFIELDS };

/* A flattened MicroModel (src/model/micro/model.jl:35-50) plus its sections
 * (src/model/micro/morphology.jl:17-29). Section indices are local to their
 * morphology; parent -1 marks a root. Parameter overrides of an insertion or
 * point process are (field, value) pairs in param_field/param_value, indexed
 * by the *_param_offsets CSR (length count+1). */
typedef struct mind_model_desc {
  double celsius;
  int32_t n_morphologies;
  const int64_t* morph_offsets; /* n_morphologies+1, into the section arrays */
  const int32_t* sec_parent;
  const double* sec_parent_x;
  const double* sec_child_x;
  const int32_t* sec_nseg;
  const double* sec_length;   /* um */
  const double* sec_diameter; /* um */
  const double* sec_Ra;       /* ohm cm */
  const double* sec_cm;       /* uF/cm2 */
  const int64_t* sec_point_offsets; /* NULL, or n_sections+1 into points (pt3d) */
  const double* points;       /* x, y, z, diameter per sample */

  int32_t n_cells;
  const int32_t* cell_morphology;
  const int64_t* cell_gid;
  const double* cell_voltage; /* initial membrane potential, mV */

  /* Density insertions in insertion order (reference insert!, model.jl:153). */
  int64_t n_density;
  const int32_t* dens_cell;
  const int32_t* dens_section;
  const int32_t* dens_mech;
  const int64_t* dens_param_offsets;

  /* Point processes in creation order (reference add_pointprocess!, model.jl:172-194);
   * cell -1 marks an artificial cell without a location. */
  int64_t n_points;
  const int32_t* pt_mech;
  const int32_t* pt_cell;
  const int32_t* pt_section;
  const double* pt_x;
  const int64_t* pt_param_offsets;

  const int32_t* param_field; /* shared pool for density and point overrides */
  const double* param_value;

  /* Spike sources (reference register_spike_source!, model.jl:286-365): kind 0 is a
   * membrane-voltage threshold detector at (cell, section, x); kind 1 is the
   * net_event output of point process `point`; kind 2 an external spike input
   * that only mind_inject_spike drives. */
  int64_t n_sources;
  const int32_t* src_kind;
  const int32_t* src_cell;
  const int32_t* src_section;
  const double* src_x;
  const double* src_threshold;
  const int32_t* src_point;
  const int64_t* src_sid;

  /* NetCons (reference connect!, model.jl:379-400) in creation order. */
  int64_t n_connections;
  const int32_t* con_source; /* source index */
  const int32_t* con_target; /* point-process index */
  const double* con_weight;  /* uS */
  const double* con_delay;   /* ms */

  /* Spike times of PatternStim points (add_pointprocess!(...; spike_times),
   * model.jl:172-194): point p's sorted times are pt_spike_times[pt_spike_offsets[p]
   * .. pt_spike_offsets[p + 1]); both NULL when no point has any. At initialize
   * each time becomes a source event of every spike source on the point
   * (reset_events!, src/runtime/utils/events.jl:92-110). */
  const int64_t* pt_spike_offsets;
  const double* pt_spike_times;

  /* Ion species (MicroModel(; ions) with IonSpecies, src/model/micro/biophysics.jl:3-25):
   * n_species entries, or 0 for default_ions (na, k, ca; src/model/micro/utils/biophysics.jl:
   * 31-42). A species a mechanism uses must be listed (with the valence its USEION VALENCE
   * declares, if any). Ion rows exist per (compartment, species) where a mechanism uses the
   * species or a setting below names it; each holds the reversal potential, the internal
   * and external concentrations, the current and its voltage derivative (model.jl:710-855). */
  int32_t n_species;
  const char* const* species_name;
  const double* species_valence;
  const double* species_internal;  /* mM */
  const double* species_external;  /* mM */
  const double* species_reversal;  /* mV */
  const int32_t* species_nernst;   /* 1: reversal from the Nernst equation */
  /* Section ion settings (add_population!(...; ions), utils/model.jl:280-320), applied in
   * order on every segment of (cell, section): what 0, 1, 2 set the internal concentration,
   * external concentration or reversal of species `species` (index into the species list)
   * to value; 3 sets its reversal policy (value 0 automatic, 1 fixed, 2 initial, 3 advance,
   * 4 dynamic); 4 sets its ion style bits (IonStyle encoding, ion_style_bits). */
  int64_t n_ion_settings;
  const int32_t* ionset_cell;
  const int32_t* ionset_section;
  const int32_t* ionset_species;
  const int32_t* ionset_what;
  const double* ionset_value;
} mind_model_desc;

/* Kernel families for the micro step. */
enum {
  MIND_METHOD_PLAIN = 0,   /* one cell per thread, node state in HBM */
  MIND_METHOD_BATCHED = 1  /* cells per workgroup, node state staged in LDS */
};

/* build! keyword arguments (src/model/build.jl:120-131) plus device choices. */
typedef struct mind_build_options {
  double dt;              /* dt_micro, ms (reference default 0.025) */
  uint64_t seed;          /* RNG seed (reference default 1) */
  int32_t method;         /* MIND_METHOD_* */
  int32_t cells_per_group;/* batched: cells per workgroup, 0 = autotune */
  int32_t window_steps;   /* steps per device window, 0 = longest the minimum delay allows */
  int64_t event_capacity; /* events per delivery-step ring slot, 0 = automatic */
  int32_t device;         /* HIP device ordinal */
  int32_t secondorder;    /* reference build! keyword (build.jl:120-131): 0 backward Euler;
                             1 and 2 Crank-Nicolson, the cable assembled with C/(dt/2) and
                             v += 2 dv (step.jl:112, cuda/micro.jl:660-679); 2 also adds
                             di/dv * dv to the ion currents before nrn_state. */
} mind_build_options;

typedef struct mind_state mind_state;

/* Record selectors (reference voltage_record / mechanism_record / ion_record,
 * recording.jl:4-76).
 * kind 0: membrane voltage at (cell, section, x);
 * kind 1: field of point process `point`;
 * kind 2: field of density mechanism `mech` at (cell, section, x);
 * kind 3: ion column `field` (MIND_ION_*) of species `mech` (index into the species list)
 *         at (cell, section, x) (ion_ref, references.jl:156-178). */
enum { MIND_ION_E = 0, MIND_ION_CI = 1, MIND_ION_CO = 2, MIND_ION_I = 3, MIND_ION_DIDV = 4 };
typedef struct mind_record {
  int32_t kind;
  int32_t cell;
  int32_t section;
  double x;
  int32_t point;
  int32_t mech;
  int32_t field;
} mind_record;

/* SimulationResult (src/runtime/recording.jl:78-87): samples at every micro step
 * boundary of the run, traces row-major [n_records][n_samples], spikes sorted by
 * (time, sid). Arrays are owned by the library until mind_result_free. */
typedef struct mind_result {
  int64_t n_samples;
  double* time;
  int64_t n_records;
  double* traces;
  int64_t n_spikes;
  double* spike_times;
  int64_t* spike_sids;
  double run_seconds; /* wall time of the step loop */
} mind_result;

/* ---- Micro-macro coupling (reference MacroModel, src/model/macro/model.jl:55-303,
 * advance_macro!, src/runtime/step.jl:156-204) ----------------------------------
 * ROIs run the Wilson-Cowan REGION mechanism (RK4 at dt_macro); ROI pairs with a
 * non-zero connectome weight run the linear delayed MACRO2MACRO projection
 * (coupling_i = sum_j W_ij E_j(t - D_ij), delays rounded to macro steps); each
 * MACRO2MICRO binding is an inhomogeneous Poisson source at rate_scale * E of
 * its ROI, rekeyed per macro step, driving one point process; each ROI's
 * MICRO2MACRO transform bins the spikes of its sources into a population rate. */
enum { WC_TAU_E, WC_TAU_I, WC_C_EE, WC_C_EI, WC_C_IE, WC_C_II, WC_A_E, WC_THETA_E, WC_A_I,
       WC_THETA_I, WC_P, WC_Q, WC_K_MICRO, WC_RATE_SCALE, WC_E0, WC_I0, WC_NPARAMS };

typedef struct mind_macro_desc {
  int32_t n_roi;
  double dt_macro;            /* ms, an integer multiple of dt_micro */
  const double* weights;      /* [n_roi * n_roi], W[target * n_roi + source] */
  const double* delays;       /* [n_roi * n_roi], ms */
  const double* wc_params;    /* [n_roi * WC_NPARAMS] */
  const double* bin_ms;       /* per ROI: MICRO2MACRO bin width, multiple of dt_macro */
  const int64_t* input_offsets;  /* n_roi + 1: MACRO2MICRO bindings of each ROI */
  const int32_t* input_point;    /* target point process of each binding */
  const double* input_weight;
  const double* input_delay;
  const int64_t* output_offsets; /* n_roi + 1: MICRO2MACRO spike sources of each ROI */
  const int32_t* output_source;  /* voltage-detector source index (its node gives mean soma v) */
} mind_macro_desc;

/* Macro result: per macro step boundary, per ROI, E, I, population rate (Hz),
 * coupling input and mean soma voltage of the ROI's sources, row-major
 * [5][n_roi][n_samples]; micro spikes as in mind_result (input spikes carry
 * the negative SIDs of their virtual sources). */
typedef struct mind_macro_result {
  int64_t n_samples;
  double* time;
  int32_t n_roi;
  double* traces;
  int64_t n_spikes;
  double* spike_times;
  int64_t* spike_sids;
  double run_seconds;
} mind_macro_result;

/* ---- Mechanisms (compile_mod / load_mod / builtin, src/compiler/frontend.jl:58-110) ----
 * A MOD text is lowered at run time into the next free mechanism slot; a model
 * that inserts it gets its cell kernels compiled for the device with hipRTC at
 * mind_build (once per set of compiled mechanisms and device, then reused).
 * Runtime mechanisms are membrane mechanisms: SUFFIX density mechanisms and
 * POINT_PROCESSes whose NET_RECEIVE (connection weight only) changes their own
 * fields, with USEION reads and writes of reversal potentials, concentrations and
 * currents through the model's ion rows; no RANDOM, net_send or net_event (those
 * are the builtin NetStim's). */
enum { MIND_KIND_DENSITY = 0, MIND_KIND_POINT = 1, MIND_KIND_ARTIFICIAL = 2 };
typedef struct mind_mech_desc {
  const char* name;      /* NEURON block name (owned by the library) */
  const char* filename;  /* as given to mind_compile_mod; "" for builtins */
  int32_t kind;          /* MIND_KIND_* */
  int32_t n_fields;      /* field columns (an array field takes one per element) */
  int32_t has_net_receive;
  int32_t has_state;     /* BREAKPOINT has SOLVE statements (nrn_state) */
  int32_t compiled;      /* 1 for a runtime slot, 0 for a builtin */
  uint64_t digest;       /* first 16 hex digits of the SHA-256 of the MOD text (0 for builtins) */
} mind_mech_desc;
/* compile_mod(source; filename) (frontend.jl:74-110): lowers `length` bytes of MOD
 * text into a mechanism slot and returns its id in *mech. Compiling the same text
 * and filename again returns the same id. Errors (MIND_EARG) name the file and
 * line in mind_last_error; MIND_EQUEUE when every slot is taken. */
int mind_compile_mod(const char* source, size_t length, const char* filename, int32_t* mech);
int mind_mech_info(int32_t mech, mind_mech_desc* out);
/* Field names (library-owned strings; element k of array field a is "a[k]",
 * 1-based) and default values of a mechanism's columns, up to `capacity`;
 * *count is the number of columns. */
int mind_mech_fields(int32_t mech, const char** names, double* defaults, int32_t capacity, int32_t* count);
/* builtin(name) / a compiled mechanism by name: the most recently compiled one of
 * that name, else the builtin. */
int mind_mech_lookup(const char* name, int32_t* mech);
/* USEION statement `index` of a mechanism (declaration order; hh: na, k): the ion name
 * (library-owned), its VALENCE (*has_valence 0 when none), and per ion column MIND_ION_E,
 * _CI, _CO, _I the field column it reads into / writes from (-1 none); reads include
 * written concentrations. *count is the number of statements. */
int mind_mech_ions(int32_t mech, int32_t index, const char** name, int32_t* has_valence, double* valence,
                   int32_t read[4], int32_t write[4], int32_t* count);
/* Per column of a compiled mechanism: whether it is a STATE variable and whether its
 * NET_RECEIVE block touches it (0 for the builtins), up to `capacity`; *count is the
 * number of columns. */
int mind_mech_columns(int32_t mech, uint8_t* is_state, uint8_t* in_receive, int32_t capacity, int32_t* count);
/* Per column: 1 for a GLOBAL, one value every instance shares (global_ref,
 * src/model/references.jl:147-154), up to `capacity`; *count is the number of
 * columns. mind_compile_mod refuses a MOD that assigns a GLOBAL (the reference runs
 * such instances one after another), so each instance's column holds the GLOBAL's
 * value. The builtins declare none. */
int mind_mech_globals(int32_t mech, uint8_t* is_global, int32_t capacity, int32_t* count);
/* C++ text of a compiled mechanism's phases (namespace mind::user, functions
 * u<id>_current/_advance/_initial/_receive over Cols), as hipRTC compiles them;
 * for host builds of the same lowering. */
int mind_mech_lowering(int32_t mech, const char** text);
/* Host only (no device): compiles the cell kernels for the compiled mechanisms
 * mechs[0..n) and GPU architecture `arch` (e.g. "gfx950") with hipRTC, as
 * mind_build does for a model that uses them, without loading the code. No
 * counterpart in the reference (a build check). */
int mind_rtc_check(const int32_t* mechs, int32_t n, const char* arch);

/* Pack and upload a model. Replaces build!/pack_micro (src/model/build.jl:120-220,
 * src/model/micro/model.jl:420-902) and allocate_simulation
 * (src/runtime/initialize.jl:4-250). */
int mind_build(const mind_model_desc* model, const mind_build_options* options,
               mind_state** out);

/* Reset clocks, queues and state and run INITIAL blocks, time-zero delivery and
 * current evaluation (src/runtime/initialize.jl:270-429). */
int mind_initialize(mind_state* state);

/* Advance to an absolute stop time that is an integer number of steps ahead
 * (src/runtime/step.jl:4-74). */
int mind_run(mind_state* state, double stop, const mind_record* records, int64_t n_records,
             int32_t record_spikes, mind_result* out);

/* Micro-macro model: the micro network plus its ROI graph (build!, build.jl:120-220,
 * with pack_macro and pack_transforms). mind_initialize also seeds the ROI state and
 * history (initialize.jl:390-424); mind_run_macro advances whole macro steps. */
int mind_build_macro(const mind_model_desc* micro, const mind_macro_desc* macro,
                     const mind_build_options* options, mind_state** out);
int mind_run_macro(mind_state* state, double stop, int32_t record_spikes, mind_macro_result* out);
void mind_macro_result_free(mind_macro_result* result);

/* One micro step (src/runtime/step.jl:77-86). */
int mind_step(mind_state* state);

void mind_result_free(mind_result* result);
void mind_state_destroy(mind_state* state);

/* Introspection used by the host layer and by tests. */
int mind_node_count(const mind_state* state, int64_t* nodes, int64_t* compartments);
int mind_read_voltage(mind_state* state, double* out, int64_t n); /* canonical node order */
int mind_read_field(mind_state* state, int32_t mech, int32_t field, double* out, int64_t n);
int mind_row_count(const mind_state* state, int32_t mech, int64_t* rows);
/* Connection events routed (ticketed, including injected ones) and gathered for
 * delivery since mind_initialize. Synchronises the state's stream. No reference
 * counterpart (a measurement read-out). */
int mind_event_counts(mind_state* state, int64_t* routed, int64_t* delivered);
int mind_time(const mind_state* state, double* t, int64_t* step);
/* read_value (src/runtime/state.jl:586-600): the current value of one record
 * selector (voltage, point-process field, density field); NetStim fields too.
 * Needs an initialized state. */
int mind_read_value(mind_state* state, const mind_record* record, double* out);
/* connection_weights (src/runtime/state.jl:603-616): the NET_RECEIVE arguments of
 * connection `connection` (0-based creation order) into out[0..capacity); *count
 * is their number (1 for the six mechanisms: the weight). */
int mind_connection_weights(const mind_state* state, int64_t connection, double* out, int32_t capacity,
                            int32_t* count);
/* inject_event! (src/runtime/state.jl:618-636): one event of connection
 * `connection` (0-based creation order) arriving at absolute time `time` >= now,
 * without its delay, ticketed at the call like any enqueued event. */
int mind_inject_event(mind_state* state, int64_t connection, double time);
/* inject_spike! (src/runtime/state.jl:638-659): a spike of external source
 * `source` (kind 2) at absolute time `time` >= now, routed over its connections
 * and recorded at its delivery call. At most 1024 injected and PatternStim spikes may fall in one
 * window. Pending injections are cleared by mind_initialize. */
int mind_inject_spike(mind_state* state, int64_t source, double time);

/* Checkpoint and restore of a micro state between runs. No counterpart in the
 * reference; they serve split runs across processes. mind_checkpoint writes the
 * state a run continues from (voltages, mechanism data, detector flags, NetStim
 * streams and self-events, clock, queued events, pending injections) into
 * buffer[0..capacity) and sets *size to its length (buffer NULL: only the size).
 * mind_restore loads one into a state built from the same model and options;
 * the next run continues bit for bit. Micro-macro states are rejected. */
int mind_checkpoint(mind_state* state, void* buffer, size_t capacity, size_t* size);
int mind_restore(mind_state* state, const void* buffer, size_t size);
int mind_window_steps(const mind_state* state, int32_t* steps);
/* Batched method: node budget per workgroup, block size, node-state layout of those
 * workgroups (0 all in LDS; 5 lean: rhs, pivots and v in LDS, the solve's node
 * records streamed from HBM; 4 lean with v in HBM) and workgroup count chosen by
 * the autotuner. */
int mind_batch_plan(const mind_state* state, int64_t* budget, int32_t* threads, int32_t* layout,
                    int64_t* groups);
/* Name of the cell kernel the batched method launches for the workgroups within
 * the budget (a static string: k_window, k_window_lean or k_window_lean_hbm, with
 * the suffix _core when every mechanism group of the model has an inlined path).
 * For measurement; no counterpart in the reference. */
int mind_batch_kernel(const mind_state* state, const char** name);
/* Replaces that plan (threads a multiple of 64 up to 1024, layout 0, 4 or 5);
 * mind_initialize must follow. */
int mind_set_batch_plan(mind_state* state, int64_t budget, int32_t threads, int32_t layout);
/* Host only (no device): builds the batched plan of a model for (budget, threads,
 * layout) and checks every index its kernels derive from the plan's tables
 * against the bounds of what it indexes. No counterpart in the reference. */
int mind_check_batch_plan(const mind_model_desc* model, double dt, int64_t budget, int32_t threads,
                          int32_t layout, int64_t* groups);
/* ---- Surrogate MLP (no counterpart in the reference) ----------------------------
 * n_in -> hidden (x layers, GELU) -> n_out on BF16 MFMA with FP32 accumulation;
 * FP32 master weights and Adam moments. Flat parameter order: per linear layer
 * W (k x n, row-major) then b (n). */
typedef struct mind_mlp mind_mlp;
int mind_mlp_create(int32_t n_in, int32_t hidden, int32_t layers, int32_t n_out, int64_t max_batch,
                    mind_mlp** out);
void mind_mlp_destroy(mind_mlp* mlp);
int mind_mlp_param_count(const mind_mlp* mlp, int64_t* n);
int mind_mlp_set_params(mind_mlp* mlp, const float* params, int64_t n);
int mind_mlp_get_params(mind_mlp* mlp, float* params, int64_t n);
int mind_mlp_forward(mind_mlp* mlp, const float* x, int64_t batch, float* y);
/* Allocation-free variant of mind_mlp_forward. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops). x (batch x n_in) and y (batch x n_out) are
 * caller-owned DEVICE pointers on the MLP's device; the activations live in a
 * static 4 MiB device array of the library (one per GPU), which limits batch to
 * 4 MiB / (8 x widest padded layer) rows (2048 at width 256; larger batches are
 * refused with MIND_EARG). stream NULL: run on the MLP's stream and return when
 * y is written; otherwise only enqueue on that hipStream_t. Calls that overlap
 * on one GPU share the array and race. Results equal mind_mlp_forward's. */
int mind_mlp_forward_noalloc(const mind_mlp* mlp, const float* x, int64_t batch, float* y, void* stream);
int mind_mlp_grad(mind_mlp* mlp, const float* x, const float* y, int64_t batch, float* loss, float* grads);
int mind_mlp_train(mind_mlp* mlp, const float* X, const float* Y, int64_t n, int64_t batch, int32_t epochs,
                   float lr, uint64_t seed, float* epoch_loss);

/* ---- Surrogate micro-macro coupling (no counterpart in the reference) -----------
 * The ROI graph of a mind_macro_desc (Wilson-Cowan RK4, delayed LinearDelay
 * coupling; its MACRO2MICRO / MICRO2MACRO bindings are ignored) integrated in one
 * persistent single-workgroup kernel with a trained MLP in place of every micro
 * population. Each ROI's drive is rate_scale * E at the start of every macro step;
 * at the end of every bin (bin_ms, the same for all ROIs) the network maps
 * [the last `history` drive bins (Hz, oldest first; the first bin repeated before
 * the run has that many), previous rate (Hz), previous mean V (mV)] to
 * [rate (Hz, clamped at 0), mean soma V (mV)] for all ROIs at once (16 per MFMA
 * tile), plus an optional linear skip map of the same normalised inputs. A bin whose raw inputs leave [x_lo, x_hi] is counted out of distribution;
 * the caller then falls back to the plain micro-macro simulation. */
typedef struct mind_surrogate_desc {
  int32_t n_in, hidden, layers, n_out; /* n_in = history + 2, n_out = 2 */
  int32_t history;
  const float* params;                 /* flat order of mind_mlp */
  int64_t n_params;
  const float* x_mean;                 /* n_in: inputs are (x - x_mean) / x_std */
  const float* x_std;
  const float* y_mean;                 /* n_out: outputs are y * y_std + y_mean */
  const float* y_std;
  const float* x_lo;                   /* n_in: training range of the raw inputs */
  const float* x_hi;
  const float* skip_w;                 /* NULL, or [n_in * n_out]: linear skip on the normalised */
  const float* skip_b;                 /* inputs added to the network output (n_out) */
  double v_rest;                       /* previous mean V before the first bin */
} mind_surrogate_desc;
typedef struct mind_surrogate mind_surrogate;
int mind_build_surrogate(const mind_macro_desc* macro, const mind_surrogate_desc* net, int32_t device,
                         mind_surrogate** out);
/* Initial ROI state and history (as mind_initialize for the plain model), rate 0, V = v_rest. */
int mind_initialize_surrogate(mind_surrogate* s);
/* Advance whole macro steps to `stop`; the result holds no spikes. ood_windows receives the
 * number of bins with out-of-range inputs and first_ood_time the end time of the first (-1 if none). */
int mind_run_surrogate(mind_surrogate* s, double stop, mind_macro_result* out, int64_t* ood_windows,
                       double* first_ood_time);
void mind_surrogate_destroy(mind_surrogate* s);

/* Host-only packing query (no device needed): node and compartment counts and,
 * when the arrays are non-NULL, the packed cable geometry in canonical node order. */
int mind_pack_nodes(const mind_model_desc* model, double dt, int64_t* nodes, int64_t* compartments,
                    int32_t* parent, double* area, double* cap, double* axial, double* axial_diag);
/* How the mechanism phases were built: the NMODL lowering of the named MOD files
 * (with their SHA-256 prefixes) or the hand lowering. No state needed. */
const char* mind_build_info(void);
/* Number of visible HIP devices (0 when none); no state needed. */
int mind_device_count(int32_t* count);
/* Measurement aid, no counterpart in the reference: read+write GB/s of a plain
 * streaming copy of `bytes` (a multiple of 32) repeated `repeats` times; the
 * known-size kernel that calibrates HBM traffic counters. */
int mind_copy_bandwidth(int32_t device, int64_t bytes, int32_t repeats, double* gbps);
/* Wait until every launch of this state has completed. */
int mind_synchronize(mind_state* state);
/* Timing of the dominant cell-kernel launch with HIP events on the state's stream
 * (batched: the launch of workgroups within the node budget; plain: the cell
 * kernel). While profiling is on, windows launch without graph replay. */
int mind_set_profiling(mind_state* state, int32_t on);
int mind_kernel_time(const mind_state* state, int64_t* launches, int64_t* steps, double* total_ms);
/* Work of that launch per step: nodes, pas rows, hh rows, Exp2Syn rows, IClamp rows, detectors. */
int mind_batch_work(const mind_state* state, int64_t* counts);
const char* mind_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
