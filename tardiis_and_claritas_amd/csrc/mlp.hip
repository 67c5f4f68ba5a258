// mlp.hip -- surrogate MLP ensemble (placeholder until the MFMA kernels land).
#include <hip/hip_runtime.h>

#include "mlp.h"

struct MlpState {
  int dummy;
};
void clr_mlp_free(MlpState* st) { delete st; }
int mlp_init(MlpState**, const clr_mlp_shape*, uint64_t, hipStream_t, std::string& err) {
  err = "surrogate kernels not built";
  return CLR_E_UNSUPPORTED;
}
size_t mlp_param_count(const MlpState*) { return 0; }
int mlp_get_params(MlpState*, float*, size_t, hipStream_t, std::string& err) { err = "no ensemble"; return CLR_E_STATE; }
int mlp_set_params(MlpState*, const float*, size_t, hipStream_t, std::string& err) { err = "no ensemble"; return CLR_E_STATE; }
int mlp_train(MlpState*, const float*, const float*, int64_t, int32_t, int32_t, float, uint64_t, float*, hipStream_t,
              std::string& err) { err = "no ensemble"; return CLR_E_STATE; }
int mlp_grad(MlpState*, const float*, const float*, int32_t, float*, float*, hipStream_t, std::string& err) {
  err = "no ensemble"; return CLR_E_STATE; }
int mlp_infer(MlpState*, const float*, int64_t, float*, float*, hipStream_t, std::string& err) { err = "no ensemble"; return CLR_E_STATE; }
int mlp_time_infer(MlpState*, int64_t, int, double*, hipStream_t, std::string& err) { err = "no ensemble"; return CLR_E_STATE; }
