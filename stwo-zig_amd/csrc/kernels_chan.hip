// Device-resident Fiat-Shamir channel step as a kernel of its own, for trees
// whose root is not produced by a parent launch (a single-leaf tree).  Parent
// launches that reach the root run the same step in the thread holding it.
#include "chan.hpp"

namespace sr {

__global__ void chan_mix_root_draw_kernel(DevChannel* __restrict__ ch, const uint32_t* __restrict__ root,
                                          uint32_t* __restrict__ root_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t r[8];
#pragma unroll
    for (int i = 0; i < 8; i++) r[i] = root[i];
    chan_mix_root_draw_thread(ch, r, root_out);
}

void launch_chan_mix_root_draw(Ctx& c, DevChannel* d_ch, const uint8_t* d_root, uint8_t* d_root_out) {
    KTimer kt(c, "channel_mix_draw", 0.0);
    hipLaunchKernelGGL(chan_mix_root_draw_kernel, dim3(1), dim3(64), 0, c.stream, d_ch, (const uint32_t*)d_root,
                       (uint32_t*)d_root_out);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
