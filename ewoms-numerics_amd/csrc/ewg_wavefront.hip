// ewg_wavefront.hip -- the tile edge buffers of the tile-wavefront ILU(0) launches
// (ewg_sweep.hip): one 64-bit word per value word, the sentinel where no value waits.
#include "ewg_wavefront.h"

namespace ewg {

namespace {

// a signalling-NaN bit pattern: no FP32 pair or FP64 value the sweeps exchange has it
constexpr unsigned long long kSentinel = 0x7FF4DEAD7FF4DEADull;

__global__ void k_wf_fill(unsigned long long *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = kSentinel;
}

}  // namespace

void wf_init_edges(const WfSync &sync, hipStream_t s) {
    hipLaunchKernelGGL(k_wf_fill, dim3(256), dim3(256), 0, s, sync.ex, sync.edge_words);
    hipLaunchKernelGGL(k_wf_fill, dim3(256), dim3(256), 0, s, sync.ey, sync.edge_words);
}

}  // namespace ewg
