// net_internal.h -- the corrector network as the solver library sees it
// (both live in libvxfea.so; the public C ABI is vxn_* in vxfea.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct vx_net;

namespace vxnet_api {
// activation workspace for B samples on an X*Y*Z node grid; work goes to st
int prepare(vx_net* n, hipStream_t st, int B, int X, int Y, int Z);
// the [B][voxel][16] bf16 input buffer of the prepared workspace
uint16_t* input(vx_net* n);
// forward from the filled input buffer; out [B][voxel][16] float32
int forward_bf16(vx_net* n, float* out);
// parameter gradient of sum(out * dout) for the last forward
int backward(vx_net* n, const float* dout, float* grad);
int adam(vx_net* n, const float* grad, float lr, float b1, float b2, float eps, int step);
long long num_params(const vx_net* n);
const char* error(const vx_net* n);
}  // namespace vxnet_api
