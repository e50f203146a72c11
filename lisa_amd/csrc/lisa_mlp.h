// The radiance-cache network shared by the device kernels and the host code:
// 12 inputs, each frequency-encoded with 6 octaves of (sin, cos) -> 144
// features (zero-padded to 160), four hidden layers of width 64 with ReLU,
// 3 linear outputs, no biases.  Weights are kept transposed per layer,
// W^T[out][in_pad], row-major, which is the layout the MFMA B operand of the
// forward pass reads as 16-byte rows.
#pragma once
#include <stdint.h>

namespace lisa {
namespace mlp {

constexpr int kInputs = 12;
constexpr int kBands = 6;
constexpr int kEnc = kInputs * kBands * 2;  // 144
constexpr int kEncPad = 160;                // K multiple of 32 for the MFMA
constexpr int kHidden = 64;
constexpr int kOut = 3;
constexpr int kOutPad = 16;
constexpr int kLayers = 5;
// per layer: (out_pad, in_pad) and offset of W^T in the packed weight array
constexpr int kOutDim[kLayers] = {kHidden, kHidden, kHidden, kHidden, kOutPad};
constexpr int kInDim[kLayers] = {kEncPad, kHidden, kHidden, kHidden, kHidden};
constexpr int kOffset[kLayers + 1] = {0, 10240, 14336, 18432, 22528, 23552};
constexpr int kParams = 23552;      // packed (padded) weights
constexpr int kRealParams = 21696;  // 144*64 + 3*64*64 + 64*3

// True (unpadded) input / output width of layer l.
constexpr int real_in(int l) { return l == 0 ? kEnc : kHidden; }
constexpr int real_out(int l) { return l == kLayers - 1 ? kOut : kHidden; }

}  // namespace mlp
}  // namespace lisa
