// BSIM3 host-side setup (model cards, size-dependent parameters, stamps).
#include "xg_bsim3.h"
#include "xg_setup.h"

namespace xg {

struct Bsim3Model {};
struct Bsim3Inst {};
struct Bsim3Plan {
  int32_t n = 0, slots = 0;
};

void bsim3_free(Circuit &c) {
  for (auto *m : c.b3_models) delete m;
  for (auto *i : c.b3) delete i;
  c.b3_models.clear();
  c.b3.clear();
}
void bsim3_plan_free(Plan &plan) {
  delete plan.b3;
  plan.b3 = nullptr;
}
int bsim3_add_model(Circuit &, int, const ParamSet &, int32_t *,
                    std::string &err) {
  err = "BSIM3 not available in this build";
  return XG_E_STATE;
}
int bsim3_add_instance(Circuit &, int, const ParamSet &, const int32_t[6],
                       int32_t, int32_t, int32_t *, std::string &err) {
  err = "BSIM3 not available in this build";
  return XG_E_STATE;
}
int bsim3_prepare(Circuit &, std::string &) { return XG_OK; }
void bsim3_pattern_entries(const Circuit &,
                           std::vector<std::array<int32_t, 2>> &) {}
int bsim3_build(Circuit &, Plan &, int64_t, std::vector<Contrib> *,
                std::vector<Contrib> *, std::string &) {
  return XG_OK;
}
int64_t bsim3_slot_count(const Plan &) { return 0; }
int bsim3_upload(const Plan &, Bsim3Dev &, UpD, UpI) { return 0; }
void bsim3_bind(Bsim3Dev &d, double *slots, int32_t *limited) {
  d.slots = slots;
  d.limited = limited;
}
int32_t bsim3_eval_slots(const Plan &) { return 0; }
int32_t bsim3_count(const Plan &) { return 0; }
int bsim3_get_offsets(const Plan &, int32_t *) { return XG_OK; }
int bsim3_get_fields(const Plan &, double *, int32_t *n) {
  *n = 0;
  return XG_OK;
}
hipError_t launch_bsim3_eval(const Bsim3Dev &d, const SolverStateDev &,
                             const LoadPtrs &, hipStream_t) {
  return d.n_slots == 0 ? hipSuccess : hipErrorNotSupported;
}

}  // namespace xg
