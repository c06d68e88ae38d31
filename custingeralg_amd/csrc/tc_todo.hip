// tc_todo.hip -- entry points not yet implemented return TC_ERR_UNSUPPORTED.
#include "tc_internal.hpp"
namespace tc {
static int unsupported(const char *f)
{
    set_last_error(TC_ERR_UNSUPPORTED, std::string(f) + " is not implemented yet");
    return TC_ERR_UNSUPPORTED;
}
int collect_timings(Graph &, int64_t, uint64_t, float *, float *, int64_t *) { return unsupported("collect_timings"); }
int mlp_train_step(float *, float *, float *, int32_t, const float *, const float *, int64_t, float,
                   float *, float *) { return unsupported("mlp_train_step"); }
int mlp_forward(const float *, const float *, int64_t, float *) { return unsupported("mlp_forward"); }
int model_load(const float *) { return unsupported("model_load"); }
int model_table(uint8_t *, int32_t *) { return unsupported("model_table"); }
int autotune(Graph &, tc_schedule *, double *) { return unsupported("autotune"); }
bool schedule_override(tc_schedule *) { return false; }
const uint8_t *learned_table_device(int32_t *dim) { *dim = 0; return nullptr; }
}  // namespace tc
