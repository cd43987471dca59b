// EPL emulator: a multilayer perceptron for the angular factor of the
// elliptical power-law deflection.
//
// Inputs (4): scaled axis ratio (q - 0.65)/0.35, scaled slope (t - 1)/0.5 and
// the unit rotated position (cos psi, sin psi), psi in the first quadrant.
// Outputs (2): Omega(psi; q, t) - (cos psi, sin psi), the part of the
// tenth-order series beyond its leading term (objects/epl.cl).
// Layers: 4 -> 128 -> 128 -> 128 -> 2 with SiLU after the hidden layers.
//
// Weights are one flat float array in this order (row-major [out][in]):
//     W1[128][4] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[2][128] b4[2]
#pragma once

namespace lsd {
namespace emu {

constexpr int kHidden = 128;
constexpr int kIn = 4;
constexpr int kOut = 2;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kHidden * kIn;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kHidden * kHidden;
constexpr int kW4 = kB3 + kHidden;
constexpr int kB4 = kW4 + kOut * kHidden;
constexpr int kParams = kB4 + kOut;

constexpr float kQ0 = 0.65f, kQS = 0.35f;  // q in [0.3, 1]
constexpr float kT0 = 1.0f, kTS = 0.5f;    // t in [0.5, 1.5]

// bf16 copies of W2 and W3 in MFMA operand order (see emulator.hip)
constexpr int kPackedWords = 2 * kHidden * kHidden / 2;  // uint32 words (2 bf16 each)

// certification grid over the training domain: kCellsQ x kCellsT cells of
// (q, t); the network builds a candidate's table only in a certified cell
constexpr int kCellsQ = 32;
constexpr int kCellsT = 32;

}  // namespace emu
}  // namespace lsd
