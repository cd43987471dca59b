// PSF factorisation for the separable convolution path.
#pragma once

#include <vector>

namespace lsd {

constexpr int kMaxRank = 16;

struct PsfFactors {
    int rank = 0;                // 0: convolve directly
    std::vector<float> factors;  // per rank: column factor [ph], row factor [pw]
};

// psf is [ph][pw] row-major
PsfFactors factor_psf(const float* psf, int pw, int ph);

}  // namespace lsd
