// Batch segment kernels for third-body binaries (templates in wave_seg.hpp).
#include "wave_seg.hpp"

namespace gbgpu {

template int launch_ll_segment_t<true>(int, int, int, int, const WaveArgs &, hipStream_t);

}  // namespace gbgpu
