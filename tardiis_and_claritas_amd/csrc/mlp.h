// mlp.h -- surrogate ensemble state (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/claritas.h"

struct MlpState;
void clr_mlp_free(MlpState* st);
int mlp_init(MlpState** st, const clr_mlp_shape* shape, uint64_t seed, hipStream_t s, std::string& err);
size_t mlp_param_count(const MlpState* st);
int mlp_get_params(MlpState* st, float* out, size_t n, hipStream_t s, std::string& err);
int mlp_set_params(MlpState* st, const float* in, size_t n, hipStream_t s, std::string& err);
int mlp_train(MlpState* st, const float* x, const float* y, const float* w, int64_t n, int32_t batch, int32_t epochs,
              float lr, uint64_t seed, float* loss, hipStream_t s, std::string& err);
int mlp_grad(MlpState* st, const float* x, const float* y, const float* w, int32_t n, float* loss, float* grads,
             hipStream_t s, std::string& err);
int mlp_infer(MlpState* st, const float* x, int64_t n, float* mean, float* stdev, hipStream_t s, std::string& err);
int mlp_time_infer(MlpState* st, int64_t n, int repeats, double* ms, hipStream_t s, std::string& err);
// Shape of the ensemble; whether the fused inference kernel serves it
// (hidden <= 256, in_dim <= 32, members * out_dim <= 80); one fused pass over
// n device-resident rows x[n][in_dim] into mean/stdev[n][out_dim].
const clr_mlp_shape* mlp_get_shape(const MlpState* st);
bool mlp_fused_ok(const MlpState* st);
int mlp_infer_device(MlpState* st, const float* x, int64_t n, float* mean, float* stdev, hipStream_t s,
                     std::string& err);
// Allocation-free inference staged in the library's static device array
// (mlp_static.hip): the array's address on `device` (hipGetSymbolAddress,
// cached per device), the largest query count one call can stage, and the
// inference itself.
int mlp_static_scratch(int device, void** out, std::string& err);
size_t mlp_static_max_rows(const MlpState* st);
int mlp_infer_static(MlpState* st, void* scratch_ptr, const float* x, int64_t n, float* mean, float* stdev,
                     hipStream_t s, std::string& err);
