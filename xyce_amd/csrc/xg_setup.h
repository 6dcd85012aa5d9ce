// Host-side setup: circuit description -> structure-of-arrays device data,
// fixed CSR pattern, per-instance stamp offsets and the ordered-reduction
// programs.  Everything here runs once per circuit.
#pragma once

#include <array>
#include <map>
#include <string>
#include <vector>

#include "xg_internal.h"

namespace xg {

struct ParamSet {
  std::map<std::string, double> v;  // upper-case name -> value (given only)
  bool given(const char *k) const { return v.count(k) != 0; }
  double get(const char *k, double dflt) const {
    auto it = v.find(k);
    return it == v.end() ? dflt : it->second;
  }
};

struct Options {
  double gmin = 1.0e-12;
  int limiter = 1;
  double temp = 300.15;
  double tnom = 300.15;
  int lu_mode = 0;  // XG_LU_AUTO, XG_LU_FRONTS or XG_LU_TASKS
};

// Processed MOSFET1 model (after processParams).
struct Mos1Model {
  int dtype = 1;
  ParamSet p;
  // resolved values
  double model_l, model_w, tnom, latDiff, jctSatCurDensity, jctSatCur, rd, rs,
      rsh, kp, cgso, cgdo, cgbo, oxideCapFactor, vt0, capBD, capBS, cj, cjsw,
      pb, mj, mjsw, fc, phi, gamma, lambda, nsub, nss, tox, uo;
  int tpg;
  bool capBDGiven, capBSGiven, cjGiven, cjswGiven, rshGiven;
  double fact1, vtnom, egfet1, pbfact1;
};

struct Mos1Inst {
  int model;
  ParamSet p;
  std::array<int32_t, 6> lids;  // d g s b d' s'  (-1 = ground)
  int32_t sta_base, sto_base;
  // derived
  double temp;
  std::array<double, M1F_N> f;
  int32_t flags;
  int card;
  double gd, gs;  // drain / source conductance
};

// A linear device as added through the C ABI.
struct LinDev {
  int32_t type;        // LIN_R / LIN_C / LIN_D / LIN_V
  int32_t lids[3];     // pos, neg, branch (-1 = ground)
  double val;          // R (ohm), C (farad) or DC value (volt)
  double pulse[8];     // has_pulse, V1, V2, TD, TR, TF, PW, PER
  int32_t sta;         // capacitor charge state slot / diode store slot
  double dpar[DIO_N];  // diode fields
};

struct ForeignStamp {
  std::vector<int32_t> gids;
  std::vector<int32_t> rows, cols;
};

// BSIM3 pieces live in xg_bsim3_setup.cpp.
struct Bsim3Model;
struct Bsim3Inst;
struct Bsim3Plan;

struct Circuit {
  int32_t n_nodes = 0, n_state = 0, n_store = 0;
  Options opt;
  std::vector<Mos1Model> mos1_models;
  std::vector<Mos1Inst> mos1;
  std::vector<Bsim3Model *> b3_models;
  std::vector<Bsim3Inst *> b3;
  std::vector<ForeignStamp> foreign;
  std::vector<LinDev> lin;
  bool have_pattern = false;
  std::vector<int32_t> rowptr, colidx;
  ~Circuit();
};

// Fixed-pattern sparse LU of each component (built by lu_analyse).
//
// Position i of component c (0 <= i < n_c) holds equation prow[i] and
// unknown qcol[i] (global ids): rows are first matched to columns so that
// every diagonal entry is structurally nonzero (a voltage source's branch
// row pairs with its node's column), then both are ordered by nested
// dissection of the symmetrised graph.  The factors live in a filled,
// structurally symmetric pattern, row-major with sorted columns; positions
// are relative to the component's block (voff).  Three task phases per
// component -- factor, forward solve, backward solve -- each a list of
// levels whose tasks are independent.  A task computes one value:
//   factor   v[dst] = (v[dst] - sum v[a] * v[b]) [/ v[div]]
//   solve    y[dst] = (y[dst] - sum v[a] * y[b]) [/ v[div]]
// with its terms in ascending order, so every result is computed by one
// thread in one fixed order.
constexpr int kLuShortOps = 4;
constexpr int kLuLongOps = 32;
struct LuPlan {
  std::vector<int32_t> prow, qcol;    // by comp_start offsets
  std::vector<int32_t> qpos;          // unknown -> position in its component
                                      // (n_nodes + 1 entries, ground -1)
  std::vector<int32_t> lu_pos;        // CSR entry -> position in its block, -1
  // the CSR entries of each component's rows in position order (ground
  // columns left out): ent_off (n_comp + 1) -> CSR entry, factor position,
  // column position; ent_rs: per component n_c + 1 row starts (relative),
  // at comp_start[c] + c
  std::vector<int32_t> ent_off, ent_k, ent_pos, ent_col, ent_rs;
  std::vector<int64_t> voff;          // n_comp + 1 value offsets
  std::vector<int32_t> ok;            // 1: matching found (static factorisation)
  std::vector<int32_t> in_lds;        // 1: factors, vector and tasks fit LDS
  std::vector<int32_t> lv_comp;       // 3 n_comp + 1: phase -> first level
  std::vector<int32_t> lv_ptr;        // level -> first task (n_levels + 1)
  std::vector<int32_t> lv_mid, lv_long;  // level -> first task with more
                                      // than kLuShortOps / kLuLongOps terms
                                      // (n_levels each); a level holds its
                                      // short tasks (one thread each), then
                                      // mid (8 lanes each), then long (one
                                      // wavefront each)
  std::vector<int32_t> t_dst, t_div;  // t_div: -1 none, -2 pivot (check only),
                                      // kLuFwdTask forward row
  std::vector<int32_t> t_op;          // task -> first op (n_tasks + 1)
  std::vector<int32_t> ops;           // pairs (a, b)
  int64_t fnnz = 0, factor_ops = 0;
  int32_t max_factor_levels = 0, max_solve_levels = 0, n_static_fail = 0;
  // components too large for task lists: supernodal multifrontal LU on
  // dense fronts (every front component together, global ids)
  std::vector<int32_t> front_comp;    // n_comp: 1 = factored by fronts
  std::vector<int32_t> fr_comp, fr_m, fr_k;  // per front
  std::vector<int64_t> fr_fo, fr_ro;  // value / row-list offsets (n_fronts + 1)
  std::vector<int32_t> fr_rows;       // front rows: global positions, ascending
  std::vector<int32_t> fr_cbmap;      // at fr_ro[s] + t (t >= k): index in parent
  std::vector<int32_t> fr_ch_ptr, fr_ch;      // children, ascending
  std::vector<int32_t> fr_lv_ptr, fr_lv;      // fronts by height (bottom-up)
  std::vector<int32_t> fr_lv_maxm;            // per level: largest front
  std::vector<int32_t> fr_ent_k;              // CSR entry -> front value
  std::vector<int64_t> fr_ent_fpos;
  int64_t fr_flops = 0;
  int32_t fr_max_m = 0, n_front_comp = 0;
};

// A device-ready plan (host copies of everything that gets uploaded).
struct Plan {
  int32_t n_nodes, n_state, n_store;
  std::vector<int32_t> rowptr, colidx;
  // MOSFET1
  int32_t m1_n = 0, m1_slots = 0;
  std::vector<int32_t> m1_inst_of, m1_wave_card, m1_lid4, m1_lid_int,
      m1_sta, m1_sto, m1_flags;
  std::vector<double> m1_cards, m1_fields;
  std::vector<int32_t> m1_offsets;  // n x 22, reference order
  std::vector<double> m1_fields_ref;  // n x M1F_N, reference order
  int64_t m1_slot_base = 0;
  // BSIM3
  Bsim3Plan *b3 = nullptr;
  // linear devices (eval order: resistors, capacitors, sources)
  int32_t lin_n = 0, lin_slots = 0;
  std::vector<int32_t> lin_type, lin_lid, lin_sta;
  std::vector<double> lin_val, lin_pulse, lin_dpar;
  int64_t lin_slot_base = 0;
  // slot buffer and reduction programs
  int64_t n_slot_values = 0;
  ReduceProgram red_vec[4];  // F, Q, dFdxdVp, dQdxdVp
  ReduceProgram red_mat[2];  // dFdx, dQdx
  int32_t n_partials = 0;    // chunk partial sums over all six programs
  int32_t limiter = 1;
  int32_t lu_mode = 0;  // Options::lu_mode
  double gmin = 1e-12;
  // connected components of the circuit graph without ground (independent
  // sub-circuits), for the per-component dense solve
  std::vector<int32_t> comp_start, comp_nodes, comp_local;
  std::vector<int64_t> comp_aoff;
  std::vector<int32_t> slot_comp;  // MOSFET1 then BSIM3 eval slots -> component
  int64_t comp_work = 0;     // doubles of dense workspace (large components)
  LuPlan lu;                 // static-pivot sparse factorisation per component
  ~Plan();
};

// One "+=" statement of the reference: out[out] += slot[src]
// (src == -1: "+= 0.0").
struct Contrib {
  int32_t out;
  int32_t src;
};

int parse_params(int32_t n, const char *const *keys, const double *vals,
                 ParamSet &out);
int mos1_model_process(Mos1Model &m, const Options &opt, std::string &err);
int mos1_instance_setup(const Mos1Model &m, Mos1Inst &in, const Options &opt,
                        std::string &err);
int build_plan(Circuit &c, Plan &plan, std::string &err);
// Diode fields from .MODEL and instance parameters (level 1, no series
// resistance); XG_E_PARAM for what the load does not model.
int diode_setup(const ParamSet &model, const ParamSet &inst, const Options &opt,
                double *fields, std::string &err);
int lu_analyse(Plan &plan, std::string &err);
// analysed components by topology (process-wide): {hits, misses, bytes};
// a new byte limit (0 = off) empties it and zeroes the counts
void lu_cache_stats(int64_t out[3]);
void lu_cache_set_limit(int64_t bytes);

// BSIM3 hooks (xg_bsim3_setup.cpp).
int bsim3_add_model(Circuit &c, int dtype, const ParamSet &p, int32_t *id,
                    std::string &err);
int bsim3_add_instance(Circuit &c, int model, const ParamSet &p,
                       const int32_t lids[6], int32_t sta, int32_t sto,
                       int32_t *id, std::string &err);
void bsim3_free(Circuit &c);
// Instance setup (size-dependent parameters, temperature), cards, checks.
int bsim3_prepare(Circuit &c, std::string &err);
// (row, col) of every BSIM3 stamp entry, ground = n_nodes.
void bsim3_pattern_entries(const Circuit &c,
                           std::vector<std::array<int32_t, 2>> &entries);
// Evaluation order, SoA data, offsets and reduction contributions.
int bsim3_build(Circuit &c, Plan &plan, int64_t slot_base,
                std::vector<Contrib> *vec, std::vector<Contrib> *mat,
                std::string &err);
int64_t bsim3_slot_count(const Plan &plan);
// For every BSIM3 eval slot, one of its non-ground nodes (-1: none/padding).
void bsim3_slot_nodes(const Plan &plan, int32_t ground, std::vector<int32_t> &out);
void bsim3_plan_free(Plan &plan);

}  // namespace xg
