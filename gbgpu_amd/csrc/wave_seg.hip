// Batch segment kernels for circular binaries (templates in wave_seg.hpp),
// and the segment dispatch of the N-bucket launches of gbgpu_wave_ll_batched.
#include "wave_seg.hpp"

namespace gbgpu {

template int launch_ll_segment_t<false>(int, int, int, int, const WaveArgs &, hipStream_t);
// instantiated in wave_seg3.hip
extern template int launch_ll_segment_t<true>(int, int, int, int, const WaveArgs &, hipStream_t);
// wave_big.hip
int launch_big_segment(int N, int precision, int pack_precision, bool third, const WaveArgs &a, hipStream_t st);

// One N bucket's segment: N <= 2048 on the uniform kernel's tiles (variant
// as for launch_ll_segment_t), N = 4096 and 8192 on the large-N tile.
// precision: arithmetic of the pass; pack_precision: layout of a.pack.
int launch_ll_segment(int N, int precision, int pack_precision, int variant, bool third, const WaveArgs &a,
                      hipStream_t st) {
    if (precision == GBGPU_MIXED && pack_precision != GBGPU_MIXED) {
        set_error("mixed segments need mixed per-binary constants");
        return GBGPU_E_ARG;
    }
    if (N == 4096 || N == 8192) return launch_big_segment(N, precision, pack_precision, third, a, st);
    return third ? launch_ll_segment_t<true>(N, precision, pack_precision, variant, a, st)
                 : launch_ll_segment_t<false>(N, precision, pack_precision, variant, a, st);
}

}  // namespace gbgpu
