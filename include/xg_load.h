/*
 * xg_load.h — C ABI of the gfx950 device-load library (libxg_load.so).
 *
 * The library evaluates Xyce MOSFET level-1 and BSIM3 (level 9/49) instances
 * on an MI355X and assembles their residual vectors and Jacobian stamps into
 * a fixed CSR pattern.  It is the drop-in target of a Device "Master"
 * (src/DeviceModelPKG/Core/include/N_DEV_Device.h:93-270): the three
 * per-Newton calls map one-to-one onto
 *
 *   Device::updateState(solVec, staVec, stoVec)            N_DEV_Device.h:214
 *       -> xg_update_state
 *   Device::loadDAEVectors(solVec, fVec, qVec, ...)        N_DEV_Device.h:248
 *       -> xg_load_vectors
 *   Device::loadDAEMatrices(dFdx, dQdx)                    N_DEV_Device.h:265
 *       -> xg_load_matrices
 *
 * Conventions (N_LAS_Builder.C:336-338, N_TIA_Assembler.C:156-162):
 *  - every vector has n_nodes + 1 entries; the last one is the ground slot
 *    (GID -1).  Writes to it happen as in the reference and are ignored by
 *    the caller;
 *  - the Jacobian is an overlap-with-ground CSR matrix with n_nodes + 1 rows,
 *    each row's columns sorted by local id with the ground column last;
 *  - outputs are accumulated with "+=": the caller zeroes them first, as
 *    the reference assembler does.  Contributions to one output entry are
 *    added in the reference's order (device type, instance insertion order,
 *    statement order within the instance), so results are bitwise
 *    reproducible run to run;
 *  - all pointers passed to the xg_update_state / xg_load_* calls are
 *    device (HBM) pointers; the xg_*_host variants take host pointers and
 *    stage through a pinned buffer (synchronous);
 *  - every call returns an int status: XG_OK (0) or a negative XG_E_* code.
 *    The library never aborts.  xg_last_error() gives a message.
 *
 * One handle serves one stream; calls on a handle are not re-entrant.
 * Work is enqueued asynchronously; xg_sync() waits for it.
 */
#ifndef XG_LOAD_H
#define XG_LOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XG_OK 0
#define XG_E_ARG (-1)       /* bad argument / unknown parameter name        */
#define XG_E_STATE (-2)     /* call out of order (e.g. before finalize)     */
#define XG_E_HIP (-3)       /* HIP runtime error                            */
#define XG_E_PATTERN (-4)   /* supplied CSR pattern lacks a stamp entry     */
#define XG_E_NOMEM (-5)     /* host or device allocation failed             */
#define XG_E_PARAM (-6)     /* parameter value the model rejects            */

typedef struct xg_circuit xg_circuit; /* setup-time builder                 */
typedef struct xg_handle xg_handle;   /* finalized, device-resident load    */

/* Per-Newton-iteration flags; mirrors the fields of SolverState that the
   MOSFET1/BSIM3 loads branch on (N_DEV_SolverState.h:75-200, set by
   DeviceMgr::setupSolverInfo_, N_DEV_DeviceMgr.C:3473-3634). */
typedef struct xg_solver_state {
  int32_t dcopFlag;
  int32_t initJctFlag;
  int32_t initFixFlag;
  int32_t initTranFlag;
  int32_t newtonIter;
  int32_t tranopFlag;
  int32_t acopFlag;
  int32_t transientFlag;
} xg_solver_state;

/* Load statistics of the last xg_update_state (device-side counts). */
typedef struct xg_stats {
  int64_t n_limited;   /* instances whose limiter fired (isConverged()=false) */
  int64_t n_nonfinite; /* instances that produced a non-finite value          */
} xg_stats;

const char *xg_last_error(void);

/* ---- setup ------------------------------------------------------------- */

/* New circuit with n_nodes non-ground unknowns.  State and store vectors
   have n_state / n_store entries; devices are given their base offsets.
   Replaces the LID/state/store registration done by Topology
   (N_TOP_CktGraphBasic.C:1092-1145, DeviceInstance::registerStateLIDs). */
int xg_circuit_new(xg_circuit **c, int32_t n_nodes, int32_t n_state,
                   int32_t n_store);
void xg_circuit_free(xg_circuit *c);

/* Device options (N_DEV_DeviceOptions.C:82-147): gmin (default 1e-12),
   voltage limiter on/off (default on), circuit temperature in kelvin
   (default 300.15) and TNOM default (300.15). */
int xg_circuit_set_options(xg_circuit *c, double gmin, int32_t limiter,
                           double temp, double tnom);

/* MOSFET level 1 model card: name/value pairs as on a .MODEL line
   (parameter table N_DEV_MOSFET1.C:117-263); dtype = +1 NMOS, -1 PMOS.
   Replaces MOSFET1::Model::Model + processParams (N_DEV_MOSFET1.C:3786-3980). */
int xg_add_mos1_model(xg_circuit *c, int32_t dtype, int32_t n,
                      const char *const *keys, const double *vals,
                      int32_t *model_id);

/* MOSFET level 1 instance (N_DEV_MOSFET1.C:76-110 parameters; constructor
   and updateTemperature N_DEV_MOSFET1.C:313-1260, 3437-3680).  lids =
   {d, g, s, b, d', s'}; d'/s' are read only when the instance has a drain /
   source resistance (else they collapse onto d / s).  -1 means ground.
   The instance owns 8 state slots at sta_base and 5 store slots at
   sto_base in the order of registerStateLIDs / registerStoreLIDs
   (N_DEV_MOSFET1.C:1486-1566). */
int xg_add_mos1_instance(xg_circuit *c, int32_t model_id, int32_t n,
                         const char *const *keys, const double *vals,
                         const int32_t lids[6], int32_t sta_base,
                         int32_t sto_base, int32_t *inst_id);

/* BSIM3 v3.2 model card (parameter table N_DEV_MOSFET_B3.C:132-1820;
   Model::processParams N_DEV_MOSFET_B3.C:8488-8596).  CAPMOD 0, 1, 2 and 3
   (default) are evaluated; any other value is XG_E_PARAM. */
int xg_add_bsim3_model(xg_circuit *c, int32_t dtype, int32_t n,
                       const char *const *keys, const double *vals,
                       int32_t *model_id);

/* count MOSFET1 / BSIM3 instances in one call, as count single calls in
   array order would: instance i takes model model_ids[i], the n_keys
   parameters keys with the values vals[i * n_keys ...], lids[6 i ...],
   sta_base[i] and sto_base[i], and its id goes to inst_ids[i].  An error stops at the failing instance
   (those before it are added), as the single calls would.  For setups of
   many instances (served batches), where one call per instance costs more
   than the setup itself.  No counterpart in the reference, which
   constructs instances one by one. */
int xg_add_mos1_instances(xg_circuit *c, const int32_t *model_ids, int32_t count,
                          int32_t n_keys,
                          const char *const *keys, const double *vals, const int32_t *lids,
                          const int32_t *sta_base, const int32_t *sto_base,
                          int32_t *inst_ids);
int xg_add_bsim3_instances(xg_circuit *c, const int32_t *model_ids, int32_t count,
                           int32_t n_keys,
                           const char *const *keys, const double *vals, const int32_t *lids,
                           const int32_t *sta_base, const int32_t *sto_base,
                           int32_t *inst_ids);

/* BSIM3 instance (N_DEV_MOSFET_B3.C:1845-2785, updateTemperature and the
   size-dependent parameters N_DEV_MOSFET_B3.C:3611-4518).  lids =
   {d, g, s, b, d', s'}; 12 state slots at sta_base and 5 store slots at
   sto_base (N_DEV_MOSFET_B3.C:2360-2364). */
int xg_add_bsim3_instance(xg_circuit *c, int32_t model_id, int32_t n,
                          const char *const *keys, const double *vals,
                          const int32_t lids[6], int32_t sta_base,
                          int32_t sto_base, int32_t *inst_id);

/* Linear devices, loaded after the MOSFETs through the same ordered
   reduction (device types in the order resistor, capacitor, source; each in
   the order added).  Node ids: -1 = ground.
   Resistor: N_DEV_Resistor.C:1073-1174 (G = 1 / r).
   Capacitor: N_DEV_Capacitor.C:1821-2060 (q = C (v+ - v-), written to state
   slot sta; no IC=).
   Voltage source: N_DEV_Vsrc.C:941-1036, branch current unknown at
   `branch`; pulse = {V1, V2, TD, TR, TF, PW, PER} (defaults resolved by the
   caller, N_DEV_SourceData.C:902-916) or NULL for a DC source; the pulse
   applies at tranop / transient iterations, dc otherwise
   (N_DEV_SourceData.C:928-1000). */
int xg_add_resistor(xg_circuit *c, int32_t pos, int32_t neg, double r);
int xg_add_capacitor(xg_circuit *c, int32_t pos, int32_t neg, double cap,
                     int32_t sta);
/* n resistors / capacitors in one call, as n calls of the above in array
   order (no counterpart in the reference: one call instead of one per
   device for large interconnect).  Stops at the first invalid device; the
   ones before it stay added, and the error message names its index. */
int xg_add_resistors(xg_circuit *c, int32_t n, const int32_t *pos,
                     const int32_t *neg, const double *r);
int xg_add_capacitors(xg_circuit *c, int32_t n, const int32_t *pos,
                      const int32_t *neg, const double *cap, const int32_t *sta);
int xg_add_vsource(xg_circuit *c, int32_t pos, int32_t neg, int32_t branch,
                   double dc, const double *pulse);

/* Stamp of a device the library does not evaluate (resistor, capacitor,
   source ...), so that the CSR pattern is the whole circuit's.
   gids[n_nodes_dev] are the device's solution ids (-1 = ground); entry e is
   (rows[e], cols[e]) in device-local numbering, as in jacobianStamp(). */
int xg_add_foreign_stamp(xg_circuit *c, int32_t n_dev_nodes,
                         const int32_t *gids, int32_t n_entries,
                         const int32_t *rows, const int32_t *cols);

/* Use a CSR pattern built elsewhere (e.g. by the reference topology) instead
   of building one.  It must contain every stamp entry. */
int xg_circuit_set_pattern(xg_circuit *c, const int32_t *rowptr,
                           const int32_t *colidx);

/* Build pattern, offsets and the ordered-reduction program; upload all
   device data.  The circuit may be freed afterwards. */
int xg_finalize(xg_circuit *c, xg_handle **h);

/* Same setup work as xg_finalize (pattern, offsets, per-instance fields,
   reduction program) without touching the GPU: the handle answers the
   xg_get_* / xg_pattern_size queries only and every load call on it fails
   with XG_E_STATE.  For checking setup on hosts without a device. */
int xg_finalize_plan(xg_circuit *c, xg_handle **h);
void xg_destroy(xg_handle *h);

/* Pattern of the finalized handle: rowptr has n_nodes + 2 entries. */
int xg_pattern_size(const xg_handle *h, int32_t *n_rows_with_ground,
                    int64_t *nnz);
int xg_get_pattern(const xg_handle *h, int32_t *rowptr, int32_t *colidx);

/* CSR positions of every instance's stamp entries, in the reference's
   registerJacLIDs order (MOSFET1: the 22 offsets of
   N_DEV_MOSFET1.C:1650-1673; BSIM3 without IC=/NQSMOD: the same 22 in the
   same order, N_DEV_MOSFET_B3.C:3150-3330).  Offsets are within the row,
   as matrixGlobalToLocal produces them. */
int xg_get_mos1_offsets(const xg_handle *h, int32_t *off /* n_inst x 22 */);
int xg_get_bsim3_offsets(const xg_handle *h, int32_t *off /* n_inst x 22 */);

/* Per-instance temperature/geometry fields the load reads, for checking the
   setup against other implementations (layout: xg_mos1_field_names). */
int xg_get_mos1_fields(const xg_handle *h, double *out, int32_t *n_fields);
int xg_get_bsim3_fields(const xg_handle *h, double *out, int32_t *n_fields);

/* ---- per Newton iteration (device pointers) ----------------------------- */

int xg_set_solver_state(xg_handle *h, const xg_solver_state *ss);

/* Master::updateState (N_DEV_MOSFET1.C:4103, N_DEV_MOSFET_B3.C:9229): reads
   x_next, sta_curr, sto_curr (and sto_next when newtonIter > 0), writes
   sta_next / sto_next (and sta_curr at initTran for BSIM3), and computes
   every stamp value of the subsequent vector and matrix loads. */
int xg_update_state(xg_handle *h, const double *x_next, double *sta_curr,
                    const double *sto_curr, double *sta_next, double *sto_next);

/* One device type's Master::updateState alone (the per-type call inside the
   loop of DeviceMgr::updateState, N_DEV_DeviceMgr.C:2258, i.e.
   DeviceMaster<T>::updateState, N_DEV_DeviceMaster.h:642): XG_DEV_MOS1,
   XG_DEV_BSIM3, or XG_DEV_LINEAR (resistors, capacitors, diodes and
   sources, evaluated by one kernel).  Calling it for every type in that
   order equals xg_update_state. */
#define XG_DEV_MOS1 1
#define XG_DEV_BSIM3 2
#define XG_DEV_LINEAR 3
int xg_update_state_type(xg_handle *h, int32_t type, const double *x_next,
                         double *sta_curr, const double *sto_curr,
                         double *sta_next, double *sto_next);

/* Master::loadDAEVectors (N_DEV_MOSFET1.C:4202, N_DEV_MOSFET_B3.C:9306). */
int xg_load_vectors(xg_handle *h, double *F, double *Q, double *dFdxdVp,
                    double *dQdxdVp);

/* Master::loadDAEMatrices (N_DEV_MOSFET1.C:4441, N_DEV_MOSFET_B3.C:9654):
   values of the two CSR matrices (same pattern). */
int xg_load_matrices(xg_handle *h, double *dFdx_vals, double *dQdx_vals);

/* All three in one enqueue (same results; fewer launches). */
int xg_load_all(xg_handle *h, const double *x_next, double *sta_curr,
                const double *sto_curr, double *sta_next, double *sto_next,
                double *F, double *Q, double *dFdxdVp, double *dQdxdVp,
                double *dFdx_vals, double *dQdx_vals);

/* xg_load_all without any allocation: the statement values and reduction
   partials go to a library-wide static device array of
   XG_NOALLOC_SCRATCH_BYTES per GPU instead of the handle's own buffers.
   No counterpart in the reference: it is for callers that cannot allocate
   device memory (code under hipGraph capture, tight serving loops,
   harnesses that forbid hipMalloc).  Results are the bits of xg_load_all.
   Limits: XG_E_ARG when the circuit's statement values and partials exceed
   the array (about 55,000 BSIM3 instances with their interconnect);
   XG_E_STATE with an active mask
   set (it keeps statement values between calls) or when the current device
   is not the handle's.  Every call on one GPU shares that GPU's array, so
   calls that overlap on one GPU (other streams, other threads) race;
   serialise them. */
#define XG_NOALLOC_SCRATCH_BYTES (32u << 20)
int xg_load_all_noalloc(xg_handle *h, const double *x_next, double *sta_curr,
                        const double *sto_curr, double *sta_next, double *sto_next,
                        double *F, double *Q, double *dFdxdVp, double *dQdxdVp,
                        double *dFdx_vals, double *dQdx_vals);

/* The three per-Newton calls with host pointers, for a caller whose
   vectors and CSR value arrays stay in host memory (the reference's
   N_LAS objects): the same arguments and results as xg_update_state,
   xg_load_vectors and xg_load_matrices, staged through a pinned buffer to
   device mirrors and back; each returns when its results are in the
   caller's arrays.  xg_update_state_host reads x_next, sta_curr, sto_curr
   and sto_next, and writes sta_next, sto_next and sta_curr (initTran);
   the loads then reduce its statement values into the caller's arrays
   (added to them, or written in XG_OUT_OVERWRITE).  The mirrors and the
   staging buffer are allocated by the first call. */
int xg_update_state_host(xg_handle *h, const double *x_next, double *sta_curr,
                         const double *sto_curr, double *sta_next, double *sto_next);
int xg_load_vectors_host(xg_handle *h, double *F, double *Q, double *dFdxdVp,
                         double *dQdxdVp);
int xg_load_matrices_host(xg_handle *h, double *dFdx_vals, double *dQdx_vals);

int xg_get_stats(xg_handle *h, xg_stats *st);
int xg_sync(xg_handle *h);

/* Stream the handle enqueues on (hipStream_t; 0 = default stream). */
int xg_set_stream(xg_handle *h, void *stream);

/* Output mode of xg_load_vectors / xg_load_matrices / xg_load_all.
   XG_OUT_ACCUMULATE (default): every contribution is added to what the
   output holds, as the reference's loads add into vectors and matrices the
   assembler zeroed (N_TIA_Assembler.C:156-162; N_DEV_DeviceMgr.C:2314-2323,
   2503-2515).  XG_OUT_OVERWRITE: every entry of the given outputs is written:
   contributions are summed from +0.0, in the same order, and entries no
   MOSFET reaches (and the ground slot, row and column) are set to +0.0 --
   the same bits as zeroing first and accumulating, without the zeroing
   pass.  No counterpart in the reference: it folds the assembler's zeroing
   into the load for callers whose other devices add afterwards. */
#define XG_OUT_ACCUMULATE 0
#define XG_OUT_OVERWRITE 1
int xg_set_output_mode(xg_handle *h, int32_t mode);

/* BSIM3 evaluation kernel selection (no counterpart in the reference: a
   launch-configuration knob for tuning and tests).  XG_EVAL_AUTO picks by
   circuit size; the others force one form.  All forms compute identical
   results. */
#define XG_EVAL_AUTO 0
#define XG_EVAL_HBM 1          /* 4 waves per block, records read from HBM  */
#define XG_EVAL_STAGED 2       /* 1 wave per block, records staged in LDS   */
#define XG_EVAL_STAGED_SPLIT 3 /* as STAGED, DC and charge parts on two waves */
int xg_set_eval_variant(xg_handle *h, int32_t variant);

/* Autotuner (no counterpart in the reference): xg_finalize / xg_upload time
   the BSIM3 eval kernel forms above, with the linear devices in trailing
   workgroups of the BSIM3 launch or in a launch of their own, on the
   circuit itself (circuits of 8 eval waves or more), keep the fastest, and
   cache the choice by circuit shape for later circuits.  All candidates
   compute the same bits.  choice[0] = the eval form, [1] = 1 if the linear
   devices share the BSIM3 launch, [2] = 0 not tuned (default choice), 1
   timed, 2 taken from the cache; us[6] = microseconds per launch of HBM,
   STAGED and STAGED_SPLIT with the linear devices shared then separate
   (NaN where not timed).  Either pointer may be NULL.  xg_set_eval_variant
   overrides the form afterwards. */
int xg_autotune_info(const xg_handle *h, int32_t choice[3], double us[6]);

/* Junction diode (N_DEV_Diode.C, level 1, RS = 0 so the internal node is
   the positive one): .MODEL parameters IS (JS), N, TT, CJO (CJ, CJ0), VJ,
   M, EG, XTI, FC, BV (VB), IBV, IRF, TNOM (Celsius), KF, AF (ignored:
   noise); instance parameters AREA, IC, TEMP (Celsius), OFF.  sto: the
   store slot of the junction voltage (li_storevd).  Loads in the Master
   loops of N_DEV_Diode.C:2007-2108 (updateState, loadDAEVectors,
   loadDAEMatrices), after capacitors and before sources.  XG_E_PARAM for
   level 2, RS != 0 or LAMBERTW != 0. */
int xg_add_diode(xg_circuit *c, int32_t pos, int32_t neg, int32_t n_model,
                 const char *const *mkeys, const double *mvals, int32_t n_inst,
                 const char *const *ikeys, const double *ivals, int32_t sto);

/* The device half of xg_finalize for a handle made by xg_finalize_plan:
   allocate and upload (host work and device work can then be done at
   different times, e.g. setup of one request while others run).  No
   counterpart in the reference. */
int xg_upload(xg_handle *h);

/* ---- Newton iteration linear algebra (the rows after the device load) */

/* Independent sub-circuits: connected components of the circuit graph with
   ground removed.  comp_of_node (n_nodes entries) may be NULL. */
int xg_components(const xg_handle *h, int32_t *n_comp, int32_t *comp_of_node);

/* Residual and Jacobian of one Newton iteration (device pointers; vectors
   n_nodes + 1, CSR values of the circuit pattern).  transient != 0:
   N_TIA_OneStep::obtainResidual / obtainJacobian (N_TIA_OneStep.C:215-330):
     order 1: rhs = -((Q - qn0)/dt + F) + dQdxdVp/dt + dFdxdVp,
              J = dQdx/dt + dFdx;
     order 2: rhs = -((Q - qn0)/dt + F/2 + fprev/2) + dQdxdVp/dt + dFdxdVp/2,
              J = dQdx/dt + dFdx/2;
   (limiter terms only when the circuit's limiter option is on).
   transient == 0 (DC operating point): rhs = -F + dFdxdVp, J = dFdx.
   The ground entry of rhs is set to 0. */
int xg_tia_assemble(xg_handle *h, int32_t transient, int32_t order, double dt,
                    const double *qn0, const double *fprev, const double *F,
                    const double *Q, const double *dFdxdVp,
                    const double *dQdxdVp, const double *dFdx,
                    const double *dQdx, double *rhs, double *J);

/* Solves J dx = rhs per component (dense LU with partial pivoting; replaces
   the KLU solve of N_LAS_AmesosSolver.C:231-470), applies x += dx, and
   writes per component stats[4c..4c+3] = {max |rhs|, weighted RMS norm of dx
   with weights reltol * max(|wa|, |wb|) + abstol (N_NLS_DampedNewton.C:
   311-335), 1 if singular else 0, size}.  active (n_comp flags, or NULL for
   all) skips converged components. */
int xg_dense_solve(xg_handle *h, const double *J, const double *rhs, double *x,
                   const double *wa, const double *wb, double reltol,
                   double abstol, const int32_t *active, double *stats);

/* The Newton linear solve per component with a static-pivot sparse LU
   (replaces the KLU refactor-and-solve of N_LAS_AmesosSolver.C:231-470).
   Every solve is checked for backward stability (finite factors, nonzero
   pivots, ||J dx - rhs||_inf <= 1e-12 (||J||_inf ||dx||_inf + ||rhs||_inf));
   a component that fails is refactored by the dense partial-pivoting LU in
   the same call.  Arguments, update and stats as xg_dense_solve. */
int xg_lu_solve(xg_handle *h, const double *J, const double *rhs, double *x,
                const double *wa, const double *wb, double reltol,
                double abstol, const int32_t *active, double *stats);

/* Sparse LU statistics: info[0] filled factor entries, [1] multiply-adds
   per factorisation, [2] most factor levels of a component, [3] most
   triangular-solve levels, [4] components without a static factorisation,
   [5] components whose factors exceed LDS, [6] components refactored with
   pivoting by the last xg_lu_solve (synchronises; -1 without a device),
   [7] factor and solve tasks.  No counterpart in the reference. */
int xg_lu_info(xg_handle *h, int64_t info[8]);

/* Analysed components by topology, shared by every circuit of the process:
   a component whose structure (rows with component-local columns in CSR
   order, source pairings, linear-solve form) was analysed before takes the
   stored ordering, fill and task lists or fronts, mapped onto its own
   nodes and entries -- the same plan the analysis gives.  Serves batches
   that repeat sub-circuits (parameter sweeps, served requests of one
   topology).  xg_lu_cache_stats: out = {hits, misses, bytes held};
   xg_lu_cache_set_limit: byte limit (default 256 MiB, 0 turns it off),
   empties the cache and zeroes the counts.  Thread-safe.  No counterpart
   in the reference, which analyses its one circuit once. */
int xg_lu_cache_stats(int64_t out[3]);
int xg_lu_cache_set_limit(int64_t bytes);

/* Fronts of more than big_m rows (default 64) are factored panel by panel
   across many workgroups (one workgroup per panel, then every column's row
   swaps and U12 rows, then the trailing update in 64 x 64 tiles); smaller
   ones by one workgroup each.  Both forms give the same bits.  No
   counterpart in the reference (tuning and tests). */
int xg_lu_front_set_big(xg_handle *h, int32_t big_m);

/* Components too large for the task lists of xg_lu_info (more than 4 M
   fill entries or 64 M multiply-adds) are factored by a supernodal
   multifrontal LU on dense fronts (partial pivoting among each front's
   fully summed rows).  info[0] such components, [1] fronts, [2] front
   levels (heights of the elimination tree), [3] doubles of front storage,
   [4] largest front order, [5] floating-point operations per
   factorisation.  No counterpart in the reference. */
int xg_lu_front_info(const xg_handle *h, int64_t info[6]);

/* Host copy of the front plan, for inspection: sizes[0..5] = fronts,
   front rows (sum of orders), children, levels, matrix entries mapped,
   components.  i32 (entries may be NULL) = fr_comp[fronts], m[fronts],
   k[fronts], rows[front rows] (global positions, ascending per front),
   cbmap[front rows] (row t >= k of a front: its index in the parent
   front), ch_ptr[fronts + 1], ch[children], lv_ptr[levels + 1],
   lv[fronts] (fronts by height), ent_k[entries] (CSR entries),
   front_comp[components]; i64 = fo[fronts + 1] (value offsets; front s
   is m x m row-major), ro[fronts + 1] (row offsets), ent_fpos[entries]
   (value position of each mapped CSR entry).  No counterpart in the
   reference. */
int xg_lu_front_export(const xg_handle *h, int64_t sizes[6], int32_t *const *i32,
                       int64_t *const *i64);

/* Linear-solve form of a circuit, set before xg_finalize: XG_LU_AUTO
   (default) factors a component by task lists when they fit in one
   workgroup's LDS and within their budgets, else by dense fronts;
   XG_LU_FRONTS factors every component by dense fronts; XG_LU_TASKS uses
   task lists wherever their budgets allow (from global memory beyond the
   LDS).  No counterpart in the reference (a solver choice, like its
   linear-solver options). */
#define XG_LU_AUTO 0
#define XG_LU_FRONTS 1
#define XG_LU_TASKS 2
int xg_circuit_set_lu_mode(xg_circuit *c, int32_t mode);

/* Host copy of the sparse LU plan, for inspection: sizes[0..5] = n_nodes,
   CSR entries, components, levels + 1, tasks, multiply-add terms; arrays
   (may be NULL, then only sizes) = prow[n_nodes], qcol[n_nodes],
   lu_pos[CSR entries], ok[components], lv_comp[3 components + 1],
   lv_ptr[levels + 1], t_dst[tasks], t_div[tasks], t_op[tasks + 1],
   ops[2 terms]; voff[components + 1] (may be NULL).  Layout as described
   for LuPlan in the library sources.  No counterpart in the reference. */
int xg_lu_export(const xg_handle *h, int64_t sizes[6], int32_t *const *arrays,
                 int64_t *voff);

/* Device-held instance state that is not in the state / store vectors
   (each diode's junction voltage of its last evaluation, which
   Master::updateState stores at the next call, N_DEV_Diode.C:2016), for
   checkpoint and restart (N_IO_RestartMgr.C:285-870 saves the device
   instances' state with the vectors).  mode 0: *n = number of doubles;
   1: copy to buf; 2: copy from buf (*n must match).  Synchronises. */
int xg_instance_state(xg_handle *h, int32_t mode, double *buf, int64_t *n);

/* Component activity for the loads: when active (device array, one flag per
   component of xg_components) is not NULL, xg_update_state skips the
   instances of components whose flag is 0 -- their statement values,
   state and store keep what their last evaluation wrote, so the loads
   reproduce that evaluation's sums.  A Newton loop over independent
   sub-circuits passes its convergence mask, so a converged sub-circuit is
   not re-evaluated while others iterate (re-evaluation would move a
   junction limiter that the iteration limit let converge).  NULL: every
   instance.  No counterpart in the reference. */
int xg_set_active_mask(xg_handle *h, const int32_t *active);

/* Step parameters from device memory (for graph-captured step loops, whose
   kernel arguments are fixed at capture): when dstep is not NULL, the
   linear sources read time = dstep[0] and breakpoint tolerance = dstep[1],
   and transient assembly reads 1/dt = dstep[2] and order = dstep[3], at
   kernel run time, instead of the values of xg_set_time / xg_tia_assemble.
   NULL restores those.  No counterpart in the reference (its time
   integrator passes them through N_TIA_TimeIntInfo on the host). */
int xg_set_step_source(xg_handle *h, const double *dstep);

/* ONESTEP predictor (N_TIA_OneStep.C:115-133): xpred = xh0 + beta1 xh1
   (+ beta2 xh2 at order 2), with xh1 = x_n - x_{n-1}, xh2 = x_{n-1} -
   x_{n-2}; beta from updateCoeffs (N_TIA_OneStep.C:1380-1405). */
int xg_tia_predict(xg_handle *h, int32_t order, double beta1, double beta2,
                   const double *xh0, const double *xh1, const double *xh2,
                   double *xpred);

/* counts[c] (device array, n_comp) = instances of component c whose
   junction voltages the last xg_update_state limited (per-component
   "all devices converged", N_NLS_DampedNewton.C:1797-1805). */
int xg_component_limited(xg_handle *h, int32_t *counts);

/* stats[4c] = max |rhs| over component c's rows (the Newton convergence
   test's residual norm, N_NLS_DampedNewton.C:1787-1810). */
int xg_rhs_norms(xg_handle *h, const double *rhs, double *stats);

/* Time (seconds) at which the next xg_update_state evaluates PULSE sources,
   and the breakpoint tolerance of the PULSE corners (SolverState::bpTol,
   N_DEV_DeviceMgr.C:3488). */
int xg_set_time(xg_handle *h, double time, double bp_tol);

/* Sizes the library was finalized with. */
int xg_sizes(const xg_handle *h, int32_t *n_nodes, int32_t *n_state,
             int32_t *n_store, int32_t *n_mos1, int32_t *n_bsim3);

/* ---- Transient driver: the DC operating point and ONESTEP steps with the
   whole Newton loop on the device (replaces the host-side loop of
   N_NLS_DampedNewton::solve, N_NLS_DampedNewton.C:400-700, and the step
   sequence of N_TIA_OneStep, with fixed steps or the reference's step
   control).  The
   iterations run as hipGraph launches (xg_lu_solve, xg_load_all,
   xg_tia_assemble, xg_rhs_norms, xg_component_limited, the non-finite
   count and the convergence test of N_NLS_DampedNewton.C:1787-1830 per
   independent sub-circuit); several iterations share one graph launch and
   the host reads the control words once per launch (xg_tran_counters),
   halving a step whose Newton solve fails.  Parameters:
   DC reltol 1e-3, abstol 1e-12, RHSTol 1e-6, deltaXTol 1, 200 iterations;
   transient 1e-2, 1e-6, 1e-2, 0.33, 20 (N_NLS_NLParams.C:108-125).

   The caller owns every vector (device pointers, as N_TIA_DataStore owns
   the reference's): lengths n1 = n_nodes + 1, nnz CSR values, n_comp
   components, n_state / n_store.  nstats (max |rhs| at 4 c) and lim are
   per-iteration accumulators: the driver clears them at xg_tran_create and
   after every Newton iteration. */
typedef struct xg_tran xg_tran;
typedef struct xg_tran_buffers {
  double *x, *x_curr, *xh1, *xh2, *xpred;  /* n1: iterate, accepted, history */
  double *sta_c, *sta_n, *sto_c, *sto_n;   /* state / store, current / next */
  double *F, *Q, *FL, *QL;                 /* n1 each: loaded vectors       */
  double *JF, *JQ, *J;                     /* nnz: dF/dx, dQ/dx, Jacobian   */
  double *rhs, *qn0, *fprev;               /* n1                            */
  double *stats, *nstats;                  /* 4 n_comp: solve / rhs norms   */
  int32_t *lim, *active, *conv;            /* n_comp                        */
  int64_t *counts;                         /* 128: non-finite count partials*/
  int32_t *ctl;                            /* 8 control words               */
  double *step;                            /* 4: time, bp_tol, 1/dt, order  */
} xg_tran_buffers;

/* Binds a driver to a finalized handle (the handle's stream and step
   source are the driver's until xg_tran_destroy).  bp_tol: PULSE corner
   tolerance. */
int xg_tran_create(xg_handle *h, const xg_tran_buffers *b, double bp_tol,
                   xg_tran **out);
/* DC operating point at t = 0 (N_ANP_Transient's tranop); on success the
   history starts there.  > 0: Newton failure (1 singular Jacobian,
   2 non-finite device values, 3 no convergence). */
int xg_tran_dcop(xg_tran *t);
/* DC operating point with continuation: mode 0 = xg_tran_dcop; 1 = gmin
   stepping (.options nonlin continuation=gmin, N_NLS_NOX_Interface.C:
   894-960): natural continuation in p from 4 to -4 with a conductance
   10^p - 10^-4 from every voltage unknown (not the sources' branch
   currents) to ground in the residual and on the Jacobian diagonal
   (N_NLS_NOX_AugmentLinSys_GStepping.C), first step -2, grown by
   1 + 0.01 ((20 - iterations) / 20)^2 after a converged solve, halved after
   a failed one (N_NLS_LOCA_StepSizeControl.C), the last step landing on -4
   (conductance 0), at most 400 solves; 2 = Newton, then gmin stepping if
   it fails (no counterpart in the reference).  Returns as xg_tran_dcop;
   xg_tran_gmin_steps gives the continuation solves of the last run. */
int xg_tran_dcop_continuation(xg_tran *t, int32_t mode);
int xg_tran_gmin_steps(xg_tran *t, int64_t *steps);
/* Start without an operating point (.TRAN ... UIC / NOOP): the solution,
   state and store start at zero, and one load there starts the charge
   and F history.  For circuits whose sources are all zero at t = 0 this
   is the DC solution itself, reached without the DC Newton solve. */
int xg_tran_uic(xg_tran *t);
/* One step of dt (order 1 for the first two steps, then 2); a step whose
   Newton solve fails is replaced by two of dt/2, recursively. */
int xg_tran_step(xg_tran *t, double dt);
/* n_steps fixed steps of dt (the steps of xg_tran_step, bit-identical
   results), probing every `every`-th: step g = first_step + s (s = 1 ..
   n_steps) with g % every == 0 writes x_curr[ids] to row g / every - 1 of
   out (device, n_ids per row) and its time to times[g / every - 1] (host,
   may be NULL).  After xg_tran_prepare_run, later steps (order 2) run in
   groups of 8 per graph launch, each step started on the device when the
   one before converged within the graph's iterations, so the host reads
   the control words once per group; a step that needs more iterations or
   fails halts the group and continues as in xg_tran_step.  No counterpart
   in the reference (its fixed-step loop is N_ANP_Transient's, one step per
   host call). */
int xg_tran_run(xg_tran *t, double dt, int64_t n_steps, const int32_t *ids, int32_t n_ids,
                int32_t every, int64_t first_step, double *out, double *times);
/* Captures xg_tran_run's graphs of 8 steps (1 to 4 iterations per step);
   call it where graph captures are safe, as at xg_tran_create.  A circuit
   with dense-front components (milliseconds and hundreds of launches per
   iteration) gets none and keeps one step per launch. */
int xg_tran_prepare_run(xg_tran *t);
/* Variable-step ONESTEP (.TRAN TSTEP TSTOP with step control), after
   xg_tran_dcop or xg_tran_uic.  The scalar control of the reference's
   N_TIA_StepErrorControl and N_TIA_OneStep: local error estimate
   ck sqrt((|x - x_pred|_w^2 + |Q - Q_n|_w^2) / 2) with the newLte weights
   (reltol max|x_n| + abstol; reltol |Q_i| + abstol), accepted when <= 1
   (N_TIA_StepErrorControl.C:668-760); completeStep / rejectStep / initialize
   (N_TIA_OneStep.C:1945-2190, 1695-1930, 1482-1640) for the next step size
   and order (1 or 2); the given breakpoint times (PULSE corners) and tstop
   as stop times at which the integration restarts.  Defaults of the
   reference: reltol 1e-3, abstol 1e-6.  xg_tran_adaptive_step takes
   attempts until one step is accepted and returns its time (tstop reached:
   no step).  xg_tran_adaptive_info: h_est = {next step, last estimate,
   stop time}, counts = {accepted, rejected by the error test, rejected by
   Newton failure, next order}. */
int xg_tran_adaptive_begin(xg_tran *t, double tstop, double tstep, const double *bps,
                           int32_t n_bp, double reltol, double abstol);
int xg_tran_adaptive_step(xg_tran *t, double *time);
int xg_tran_adaptive_info(xg_tran *t, double h_est[3], int64_t counts[4]);
/* Eviction for batches of independent circuits (no counterpart in the
   reference, which runs one circuit): with on = 1, a step whose Newton
   solve still fails after halving down to min_dt (default 1e-18) takes the
   components that did not converge out of the solve -- their unknowns stay
   at the last accepted values -- and is taken again for the others.
   xg_tran_evicted: flags[c] = 1 for evicted components (n_comp host
   entries, may be NULL), *n = how many. */
int xg_tran_set_eviction(xg_tran *t, int32_t on, double min_dt);
int xg_tran_evicted(xg_tran *t, int32_t *flags, int64_t *n);
/* out[k] = x_curr[ids[k]] on the driver's stream (device arrays).  An id
   outside 0 .. n_nodes - 1 (ground included) reads nothing and gives NaN. */
int xg_tran_probe(xg_tran *t, const int32_t *ids, int32_t n, double *out);
/* Time of the accepted solution; counters[0..3] = steps, Newton
   iterations, failed steps, last failure word. */
int xg_tran_info(xg_tran *t, double *time, int64_t counters[4]);
/* Restart: the accepted time, step count and counters of a checkpoint
   (the caller restores the vectors it owns). */
int xg_tran_restore(xg_tran *t, double time, int64_t n_steps, int64_t newton_total,
                    int64_t step_failures);
/* Driver counters: out[0] control-word reads (host round trips), [1]
   steps, [2] Newton iterations, [3] iterations the next step's graph
   holds.  A step's graph holds as many iterations as the previous step
   took (at most 4), every launch of an iteration after convergence doing
   nothing, so a step usually costs one launch and one read.  No
   counterpart in the reference. */
int xg_tran_counters(xg_tran *t, int64_t out[4]);
/* 0: launch every iteration's kernels one by one instead of a graph. */
int xg_tran_set_graphs(xg_tran *t, int32_t on);
int xg_tran_sync(xg_tran *t);
int xg_tran_destroy(xg_tran *t);

#ifdef __cplusplus
}
#endif
#endif
