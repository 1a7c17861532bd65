// rtg_wave.hip — wavefront engine (per-depth kernels over compacted queues).
#include "rtg_internal.hpp"

namespace rtg {

int launch_wavefront(const SceneRes &, const rtg_params &, int, int, const RenderOut &, hipStream_t, Arena &) {
    return fail(RTG_E_INVALID, "wavefront engine not built yet");
}

}  // namespace rtg
