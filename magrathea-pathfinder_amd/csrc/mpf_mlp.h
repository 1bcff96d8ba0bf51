// Internal interface of the learned first-guess network (mpf_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/mpf.h"
#include "mpf_device.h"

namespace mpf {

struct MlpModel;

bool mlp_create(MlpModel** out, hipStream_t st, uint64_t seed, std::string& err);
void mlp_destroy(MlpModel* M);
// get: MPF_MLP_NSTATE floats out; set: MPF_MLP_NSTATE floats in (either may be null)
bool mlp_state(MlpModel* M, float* get, const float* set, std::string& err);
bool mlp_train(MlpModel* M, const float* feat, const float* target, uint64_t n, uint32_t epochs, uint32_t batch,
               float lr, uint64_t seed, float* loss_per_epoch, std::string& err);
bool mlp_gradient(MlpModel* M, const float* feat, const float* target, uint64_t n, float* grad, float* loss,
                  std::string& err);
bool mlp_infer_dev(MlpModel* M, const float* dfeat, uint64_t n, float* dout, std::string& err);
float mlp_last_ms(const MlpModel* M);   // duration of the last inference kernel (ms), -1 if none

// order (may be null): the source each thread takes (launch_source_order)
hipError_t launch_features(const DevTree& t, const double* src, uint64_t n, const uint32_t* order, double length,
                           float* feat, hipStream_t st);
hipError_t launch_seed(const double* src, const float* feat, const float* pred, uint64_t n, double* init,
                       hipStream_t st);
hipError_t launch_mfma_probe(const float* a, const float* b, float* c, hipStream_t st);
hipError_t launch_mfma_probe32(const float* a, const float* b, float* c, int split, hipStream_t st);

}  // namespace mpf
