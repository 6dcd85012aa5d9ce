// Internal declarations shared by the host setup (xg_setup.cpp), the C ABI
// (xg_api.hip) and the kernels (xg_mos1.hip, xg_bsim3.hip, xg_reduce.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/xg_load.h"

namespace xg {

// ---------------------------------------------------------------------------
// MOSFET level 1.
//
// Card: quantities that depend only on the model and the temperature; one
// card per (model, temperature) pair.  Staged in LDS by the eval kernel.
enum Mos1Card : int {
  M1C_DTYPE = 0, M1C_GAMMA, M1C_LAMBDA, M1C_MJ, M1C_MJSW, M1C_TPHI, M1C_TVBI,
  M1C_TVTO, M1C_VT, M1C_TDEPCAP, M1C_TBULKPOT, M1C_GMIN, M1C_LIMITER,
  M1C_MJ_EQ_HALF,    // 1.0 when MJ == 0.5 (sqrt fast path)
  M1C_MJSW_EQ_HALF,  // 1.0 when MJSW == 0.5
  M1C_MJ_EQ_MJSW,    // 1.0 when MJ == MJSW
  M1C_N = 16
};

// Per-instance fields (structure of arrays, one array per field, indexed by
// evaluation slot).  These are the reference instance members that
// updateTemperature / the constructor leave behind and the load reads.
enum Mos1Field : int {
  M1F_BETA = 0, M1F_DSATCUR, M1F_SSATCUR, M1F_SVCRIT, M1F_DVCRIT, M1F_CBS,
  M1F_CBSSW, M1F_CBD, M1F_CBDSW, M1F_F2S, M1F_F3S, M1F_F4S, M1F_F2D, M1F_F3D,
  M1F_F4D, M1F_OXIDECAP, M1F_GSO, M1F_GDO, M1F_GBO, M1F_GD, M1F_GS, M1F_M,
  M1F_ICVDS, M1F_ICVGS, M1F_ICVBS,
  M1F_N = 25
};

// Per-instance flag bits.
enum : int32_t { XF_OFF = 1, XF_IC = 2, XF_HAS_DP = 4, XF_HAS_SP = 8 };

// Stamp-value slots written per MOSFET1 instance, in the order the reference
// Master executes its "+=" statements.  F/Q/limiter entries first, then the
// 17 dF/dx and 14 dQ/dx entries (N_DEV_MOSFET1.C:4202-4430, 4441-4722).
enum Mos1Slot : int {
  M1S_F_D = 0, M1S_F_G, M1S_F_S, M1S_F_B, M1S_F_DP, M1S_F_SP,
  M1S_Q_G, M1S_Q_B, M1S_Q_DP, M1S_Q_SP,
  M1S_FL_B, M1S_FL_DP, M1S_FL_SP,
  M1S_QL_G, M1S_QL_B, M1S_QL_DP, M1S_QL_SP,
  M1S_JF0,          // 17 dF/dx entries
  M1S_JQ0 = M1S_JF0 + 17,  // 14 dQ/dx entries
  M1S_N = M1S_JQ0 + 14
};

// ---------------------------------------------------------------------------
// Ordered reduction program for one output array.  The sources of each
// output entry are slot indices in the reference's accumulation order; -1
// means "+ 0.0" (a structurally zero reference statement whose only effect
// is to turn -0.0 into +0.0) and kPad means "+ -0.0", the exact identity,
// used to pad fixed-width records.
//
//  * short entries: at most four sources, record i is the fixed-width list
//    ssrc4[i * 4 .. +4) for entry sdst4[i]; five to kShort sources, the list
//    ssrc[i * kShort .. +kShort) for entry sdst[i];
//  * the other entries' sources are stored contiguously in src and cut into
//    chunks of at most kChunk.  Chunk c covers src[cptr[c] .. cptr[c+1]); if
//    it is the whole segment of entry e then cdst[c] = e, else -1 and its
//    sum goes to partial[pbase + c].  Multi-chunk entry i (mtgt[i]) owns
//    chunks [mchunk[2i], mchunk[2i+1]);
//  * zidx lists the entries no source reaches (and the ground slot, row and
//    column), which the overwrite output mode sets to +0.0.
constexpr int32_t kShort = 8;
constexpr int32_t kChunk = 32;
constexpr int32_t kPad = -2;

struct ReduceProgram {
  std::vector<int32_t> ssrc4, sdst4;
  std::vector<int32_t> ssrc, sdst;
  std::vector<int32_t> src;
  std::vector<int32_t> cptr, cdst;
  std::vector<int32_t> mtgt, mchunk;
  std::vector<int32_t> zidx;
  int32_t pbase = 0;
};

// Device-side view of a reduction program.
struct ReduceDev {
  int32_t n_short4 = 0;
  int32_t n_short = 0;
  int32_t n_chunk = 0;
  int32_t n_multi = 0;
  int32_t n_zero = 0;
  int32_t pbase = 0;
  int32_t *ssrc4 = nullptr;
  int32_t *sdst4 = nullptr;
  int32_t *ssrc = nullptr;
  int32_t *sdst = nullptr;
  int32_t *src = nullptr;
  int32_t *cptr = nullptr;
  int32_t *cdst = nullptr;
  int32_t *mtgt = nullptr;
  int32_t *mchunk = nullptr;
  int32_t *zidx = nullptr;
};

// Linear devices (resistor, capacitor, independent voltage source), loaded
// through the same statement / ordered-reduction path as the MOSFETs.
enum LinType : int32_t { LIN_R = 0, LIN_C = 1, LIN_D = 2, LIN_V = 3 };
constexpr int kLinStatements = 20;  // the diode's; the others use fewer

// Per-instance diode fields (temperature- and area-scaled, N_DEV_Diode.C
// updateTemperature :1332-1470 and Model::processParams :1772-1793).
enum DioField : int32_t {
  DIO_ISAT = 0,   // tSatCur * AREA
  DIO_VTE,        // N * k T / q
  DIO_VCRIT,      // tVcrit
  DIO_IRF,        // tIRF
  DIO_BRKV,       // tBrkdwnV (0: no breakdown)
  DIO_BVGIVEN,    // 1 if BV was given
  DIO_BV,         // BV
  DIO_TJCAP,      // tJctCap
  DIO_CZERO,      // tJctCap * AREA
  DIO_TJPOT,      // tJctPot
  DIO_TDEPCAP,    // tDepCap = FC * tJctPot
  DIO_TF1,        // tF1
  DIO_M,          // grading coefficient
  DIO_F2,
  DIO_F3,
  DIO_TT,         // transit time
  DIO_GSPR,       // tCOND * AREA (0: no series resistance)
  DIO_OFF,        // OFF flag
  DIO_HASIC,      // IC given
  DIO_IC,         // IC value
  DIO_N
};

struct LinArgs {
  int32_t n_slots = 0;            // padded (multiple of 64), grouped by type
  const int32_t *type = nullptr;  // n_slots; -1 = padding
  const int32_t *lid = nullptr;   // 3 x n_slots: pos, neg, branch (ground = n_nodes)
  const double *val = nullptr;    // n_slots: G (R), C (C), DC value (V)
  const double *pulse = nullptr;  // 8 x n_slots: has_pulse, V1, V2, TD, TR, TF, PW, PER
  const int32_t *sta = nullptr;   // n_slots: capacitor charge state slot,
                                  // diode store slot (Vd)
  const double *dpar = nullptr;   // DIO_N x n_slots diode fields
  double *dvd = nullptr;          // n_slots: each diode's last junction voltage
  int32_t *limited = nullptr;     // n_slots: limiter / non-finite flags
  double gmin = 1e-12;
  int32_t limiter = 1;
  double *slots = nullptr;        // kLinStatements x n_slots (in the shared buffer)
};

struct SolverStateDev {
  int32_t dcop, initJct, initFix, initTran, newtonIter, tranop, acop,
      transient;
};

// Arguments of the MOSFET1 fused eval kernel.
struct Mos1Args {
  int32_t n_slots;          // padded evaluation slots (multiple of 64)
  const int32_t *inst_of;   // eval slot -> reference instance index, -1 pad
  const int32_t *wave_card; // wave index -> card index
  const double *cards;      // n_cards x M1C_N
  const double *fields;     // M1F_N x n_slots
  const int32_t *lid4;      // 4 x n_slots (d, g, s, b)
  const int32_t *lid_int;   // 2 x n_slots (d', s')
  const int32_t *sta_base;  // n_slots
  const int32_t *sto_base;  // n_slots
  const int32_t *flags;     // n_slots
  double *slots;            // M1S_N x n_slots
  int32_t *limited;         // n_slots (0/1)
};

struct LoadPtrs {
  const double *x;
  double *sta_curr;
  const double *sto_curr;
  double *sta_next;
  double *sto_next;
  // optional: component activity (xg_set_active_mask) and the component of
  // each eval slot of the kernel's device type; instances of inactive
  // components are not evaluated (their statement values, state and store
  // keep what the last evaluation wrote)
  const int32_t *active = nullptr;
  const int32_t *comp = nullptr;
  // optional: a device stop word; when it holds nonzero the launch does
  // nothing (iterations after convergence inside a multi-iteration graph)
  const int32_t *gate = nullptr;
};

__device__ __forceinline__ bool skip_slot(const LoadPtrs &p, int slot) {
  if (!p.active) return false;
  const int32_t c = p.comp[slot];
  return c >= 0 && !p.active[c];
}

// Kernel launchers (defined in the .hip files).
hipError_t launch_mos1_eval(const Mos1Args &a, const SolverStateDev &ss,
                            const LoadPtrs &p, hipStream_t s);
// overwrite = false: outputs accumulate ("+=", the reference's loads into
// zeroed vectors); true: every entry is written, contributions summed from
// +0.0 and entries without contributions set to +0.0.
hipError_t launch_reduce(int n_out, const ReduceDev *progs, double *const *outs,
                         const double *slots, double *partial, bool overwrite,
                         hipStream_t s, const int32_t *gate = nullptr);
// Newton-iteration assembly (N_TIA_OneStep.C:215-330).
struct TiaAssembleArgs {
  int64_t n1 = 0, nnz = 0;
  int32_t transient = 0, order = 1, limiter = 1;
  double inv_dt = 0.0;
  const double *dstep = nullptr;  // if set: 1/dt = dstep[2], order = dstep[3]
  const double *qn0, *fprev, *F, *Q, *FL, *QL, *dFdx, *dQdx;
  double *rhs, *J;
  // optional (the native Newton loop): max |rhs| per component folded into
  // norms[4 c] (which must hold +0.0) by an exact atomic max; NaN entries
  // are skipped, as fmax skips them
  const int32_t *node_comp = nullptr;
  double *norms = nullptr;
  const int32_t *gate = nullptr;  // stop word (LoadPtrs::gate)
};

// Dense LU solve per connected component (ground excluded).
struct DenseSolveArgs {
  int32_t n_comp = 0;
  int32_t any_large = 0;          // some component exceeds the LDS size
  const int32_t *cstart;          // n_comp + 1: component -> first node
  const int32_t *nodes;           // nodes grouped by component
  const int32_t *local;           // node -> index within its component (-1 ground)
  const int64_t *aoff;            // component -> offset of its matrix in work
  const int32_t *rowptr, *colidx; // circuit CSR (ground last)
  const double *J, *rhs;
  double *x;                      // Newton iterate, updated in place
  const double *wa, *wb;          // weights: reltol * max(|wa|, |wb|) + abstol
  const int32_t *gate = nullptr;  // stop word (LoadPtrs::gate)
  double reltol, abstol;
  const int32_t *active;          // n_comp flags or null (all)
  int32_t norm_only = 0;          // only stats[0] = max |rhs| per component
  int32_t only_flagged = 0;       // inactive components keep their stats
  double *work, *bwork;           // dense matrices / right-hand sides
  double *stats;                  // n_comp x 4: max|rhs|, wRMS(dx), singular, n
};

// t_div of a forward-substitution row (y(i) = b(i) - sum L(i, k) y(k))
// scheduled among the factor tasks: its value goes to the vector, its terms
// read L from the factors and y from the vector.
constexpr int32_t kLuFwdTask = -3;

// Static-pivot sparse LU per component (task lists of lu_analyse).
struct SparseLuArgs {
  int32_t n_comp = 0;
  int32_t any_global = 0;         // some component's factors exceed LDS
  const int32_t *cstart;          // n_comp + 1
  const int32_t *prow, *qcol;     // position -> global equation / unknown
  const int32_t *lu_pos;          // CSR entry -> position in its block (-1)
  const int64_t *voff;            // n_comp + 1 value offsets
  const int32_t *ok;              // static factorisation available
  const int32_t *in_lds;          // factors, vector and tasks fit in LDS
  const int32_t *lv_comp, *lv_ptr, *lv_mid, *lv_long;
  const int4 *tasks;              // {dst, div, first op, end op}
  const int2 *ops;
  const int32_t *ent_off, *ent_k, *ent_pos, *ent_col, *ent_rs;
  double *Jp, *bp;                // J in entry order, rhs in position order
  int32_t n_ent = 0, n_nodes = 0;
  const double *J, *rhs;
  double *x;
  const double *wa, *wb;
  double reltol, abstol;
  double max_berr;                // largest accepted normwise backward error
  const int32_t *active;
  double *val_work, *vec_work;    // factors / solution of global components
  double *stats;
  int32_t *fallback;              // n_comp: 1 = refactor with pivoting
  const int32_t *front = nullptr; // n_comp: 1 = factored by dense fronts
  const int32_t *gate = nullptr;  // stop word (LoadPtrs::gate)
};

// Dense fronts (multifrontal LU) of the components too large for task
// lists; the plan is LuPlan's fr_* arrays (xg_setup.h).
constexpr int kFrontMaxM = 3481;  // largest front a workgroup's panel holds
// Fronts of more rows are factored by many workgroups per panel.
constexpr int32_t kFrontBigM = 64;
struct FrontArgs {
  int32_t n_fronts = 0, n_comp = 0, n_levels = 0;
  int64_t n_vals = 0, n_ent = 0;
  const int32_t *comp, *m, *k, *rows, *cbmap, *ch_ptr, *ch, *lv, *ent_k, *front_comp;
  const int64_t *fo, *ro, *ent_fpos;
  double *vals, *wrhs, *v;        // fronts, front right-hand sides, positions
  int32_t *piv, *bad;             // row swaps (at ro), failed components
  const int32_t *active;
  const int32_t *gate = nullptr;  // stop word (LoadPtrs::gate)
  const int32_t *lv_ptr_host, *lv_maxm_host;  // host arrays: the launch loop
  // host copies for the many-workgroup form of fronts of more than big_m
  // rows
  const int32_t *lv_host, *m_host, *k_host, *ch_ptr_host, *ch_host;
  int32_t big_m = kFrontBigM;
  // those fronts grouped by level (device), with the host's level offsets
  const int32_t *big = nullptr, *big_host = nullptr, *big_ptr_host = nullptr;
  // components of more than kFinishSplit unknowns finish in chunks of
  // kFinishChunk rows: per component split[c] (1 = chunked), the chunks'
  // component and first row, the chunked components and their chunk
  // ranges, per-chunk partials (4 doubles) and the per-component verdict
  // children's contribution rows by target: front s's rows that receive
  // contributions are xr_row[xr_off[s], xr_off[s + 1]); target q takes
  // child rows (ext_ch, ext_i)[ext_ptr[q], ext_ptr[q + 1]) in child order
  const int64_t *xr_off = nullptr, *ext_ptr = nullptr, *xr_off_host = nullptr;
  const int32_t *xr_row = nullptr, *ext_ch = nullptr, *ext_i = nullptr;
  // rows of more than kLongRow Jacobian entries (supply rails) per front
  // component, checked by a whole workgroup: long_rows[long_ptr[c], ..)
  const int32_t *long_ptr = nullptr, *long_rows = nullptr;
  int32_t n_fin_chunks = 0, n_fin_comps = 0;
  const int32_t *fin_split = nullptr, *fin_chunk_comp = nullptr, *fin_chunk_beg = nullptr;
  const int32_t *fin_comps = nullptr, *fin_cptr = nullptr;
  double *fin_part = nullptr;
  int32_t *fin_good = nullptr;
};
constexpr int32_t kFinishSplit = 8192;
constexpr int32_t kLongRow = 256;
constexpr int32_t kFinishChunk = 2048;
int front_form(int32_t max_m);
hipError_t launch_fronts(const FrontArgs &f, const SparseLuArgs &a, hipStream_t s);

hipError_t launch_tia_assemble(const TiaAssembleArgs &a, hipStream_t s);
hipError_t launch_dense_solve(const DenseSolveArgs &a, hipStream_t s);
int dense_lds_max();
constexpr int kDenseMax = 2048;  // largest component with a dense form
hipError_t launch_sparse_lu(const SparseLuArgs &a, hipStream_t s);
// LDS bytes a component needs for the in-LDS form, and the budget
int64_t sparse_lu_lds_bytes(int64_t n, int64_t fnnz, int64_t tasks, int64_t ops,
                            int64_t levels);
int64_t sparse_lu_lds_budget();
hipError_t launch_predict(int64_t n1, int32_t order, double beta1, double beta2,
                          const double *xh0, const double *xh1, const double *xh2,
                          double *out, hipStream_t s);

hipError_t launch_linear_eval(const LinArgs &a, const SolverStateDev &ss,
                              const LoadPtrs &p, double time, double bp_tol,
                              const double *dstep, hipStream_t s);
hipError_t launch_component_limited(const int32_t *flags, const int32_t *slot_comp,
                                    int32_t n, int32_t *cnt, int32_t n_comp,
                                    hipStream_t s);
hipError_t launch_iteration_counts(const int32_t *flags, const int32_t *slot_comp,
                                   int32_t n, int32_t *comp_cnt, int64_t *out,
                                   hipStream_t s, const int32_t *gate = nullptr);
hipError_t launch_count_flags(const int32_t *flags, int32_t n, int64_t *out,
                              hipStream_t s);

// Handle internals for the transient driver (xg_tran.hip).
hipError_t handle_count_flags(xg_handle *h, int64_t *out128);
// The native Newton loop's bookkeeping of one iteration in two launches:
// assembly with the per-component max |rhs| (norms, zeroed beforehand),
// then one pass over the instance flags for the per-component limiting
// counts (comp_cnt, zeroed beforehand) and the 64-block limited /
// non-finite counts (out128).
int handle_assemble_norms(xg_handle *h, int32_t transient, const double *qn0,
                          const double *fprev, const double *F, const double *Q,
                          const double *FL, const double *QL, const double *JF,
                          const double *JQ, double *rhs, double *J, double *norms);
hipError_t handle_iteration_counts(xg_handle *h, int32_t *comp_cnt, int64_t *out128);
hipStream_t handle_stream(const xg_handle *h);
// The stop word the launches of the next enqueues check (nullptr: none).
void handle_set_gate(xg_handle *h, const int32_t *gate);
int32_t handle_n_comp(const xg_handle *h);
int64_t handle_n_nodes(const xg_handle *h);
int32_t handle_n_fronts(const xg_handle *h);
// Voltage unknowns (every unknown but the voltage sources' branch currents)
// with a diagonal Jacobian entry, and that entry's CSR position.
void handle_voltage_diagonal(const xg_handle *h, std::vector<int32_t> &ids,
                             std::vector<int32_t> &pos);
void handle_sizes(const xg_handle *h, int64_t *n_state, int64_t *n_store, int64_t *nnz);

}  // namespace xg
