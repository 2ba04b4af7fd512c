// mlp.hpp — sizes and parameter layout of the warm-start network
// (12 -> 128 -> 128 -> 128 -> 3, GELU), shared by the device code, the host
// glue and the C ABI.
#pragma once

#include <cstdint>

namespace cvmlp {
constexpr int IN = 12, H = 128, OUT = 3;
// FP32 master vector: W0 [IN x H] (row-major, input-major), b0 [H], W1, b1,
// W2, b2 [H x H, H], W3 [H x OUT], b3 [OUT].
constexpr int oW0 = 0, ob0 = IN * H, oW1 = ob0 + H, ob1 = oW1 + H * H, oW2 = ob1 + H, ob2 = oW2 + H * H,
              oW3 = ob2 + H, ob3 = oW3 + H * OUT, NPARAM = ob3 + OUT;
static_assert(NPARAM == 35075, "parameter count");
// BF16 operand tiles (32 x 16, 512 values): forward W_l, transposed W_1..W_3.
constexpr uint32_t FW_OFS[5] = {0, 1 * 8 * 512, 1 * 8 * 512 + 4 * 8 * 512, 1 * 8 * 512 + 8 * 8 * 512,
                                1 * 8 * 512 + 8 * 8 * 512 + 4 * 1 * 512};
constexpr uint32_t FW_TOTAL = FW_OFS[4];
constexpr uint32_t BW_OFS[5] = {0, 0, 4 * 8 * 512, 8 * 8 * 512, 8 * 8 * 512 + 1 * 8 * 512};
constexpr uint32_t BW_TOTAL = BW_OFS[4];
}  // namespace cvmlp
