// Host build of pack_frame_runs (clifft_amd/csrc/svm_host.cpp) for the CPU
// test of the frame superinstructions.  The test writes the superinstruction
// constants (from svm_device.h) and the function's text to frame_pack.inc.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "clifft_svm.h"

#include "frame_pack.inc"

extern "C" void pack(clifft_instruction* code, uint64_t n, const uint8_t* idle) {
    const std::vector<uint8_t> v(idle, idle + n);
    pack_frame_runs(code, n, v);
}
