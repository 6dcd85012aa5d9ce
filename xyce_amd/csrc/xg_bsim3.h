// BSIM3 device-side data and the hooks the C ABI uses.
#pragma once

#include <functional>
#include <vector>

#include "xg_internal.h"

namespace xg {

struct Plan;

struct Bsim3Dev {
  int32_t n_slots = 0;              // padded evaluation slots (multiple of 64)
  int32_t n_sdp = 0;                // size-dependent records
  const int32_t *inst_of = nullptr; // eval slot -> instance, -1 = padding
  const int32_t *wave_card = nullptr;
  const double *cards = nullptr;    // n_cards x B3C_N
  const double *sdp = nullptr;      // B3R_N x n_sdp
  const double *geo = nullptr;      // B3I_N x n_slots
  const int32_t *lid4 = nullptr;    // 4 x n_slots
  const int32_t *lid_int = nullptr; // 2 x n_slots
  const int32_t *sta_base = nullptr;
  const int32_t *sto_base = nullptr;
  const int32_t *flags = nullptr;
  const int32_t *sdp_idx = nullptr; // n_slots
  // Per-wave blocks of the slots' size-dependent + junction fields
  // ((n_slots / 64) x (B3R_N + B3I_N) x 64), for the LDS-staged kernel.
  const double *rec = nullptr;
  int32_t variant = 0;  // XG_EVAL_* of include/xg_load.h
  double *slots = nullptr;          // B3S_N x n_slots (inside the shared buffer)
  int32_t *limited = nullptr;
  int64_t slot_base = 0;
};

using UpD = std::function<double *(const std::vector<double> &)>;
using UpI = std::function<int32_t *(const std::vector<int32_t> &)>;

int bsim3_upload(const Plan &P, Bsim3Dev &d, UpD upd, UpI upi);
void bsim3_bind(Bsim3Dev &d, double *slots, int32_t *limited);
int32_t bsim3_eval_slots(const Plan &P);
int32_t bsim3_count(const Plan &P);
int bsim3_get_offsets(const Plan &P, int32_t *off);
int bsim3_get_fields(const Plan &P, double *out, int32_t *n_fields);
// Linear devices evaluated by trailing workgroups of the BSIM3 eval launch
// (the two device types have no data in common, so their evaluations run
// side by side on the GPU instead of one after the other).
struct LinTail {
  LinArgs a;
  LoadPtrs p;
  double time = 0.0, bp_tol = 0.0;
  const double *dstep = nullptr;
};
// tail: null for the BSIM3 instances alone.
hipError_t launch_bsim3_eval(const Bsim3Dev &d, const SolverStateDev &ss,
                             const LoadPtrs &p, hipStream_t s,
                             const LinTail *tail = nullptr);

}  // namespace xg
