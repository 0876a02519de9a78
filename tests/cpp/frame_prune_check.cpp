// Host build of identity_frame_ops (clifft_amd/csrc/svm_host.cpp) for the
// CPU test of the frame pruning.  The test writes the function's text, as it
// stands in svm_host.cpp (or with a planted mistake), to frame_prune.inc.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "clifft_svm.h"

#include "frame_prune.inc"

extern "C" void idle_mask(const clifft_program_view* v, uint8_t* out) {
    const std::vector<uint8_t> m = identity_frame_ops(v);
    if (!m.empty()) std::memcpy(out, m.data(), m.size());
}
