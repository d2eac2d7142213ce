// Initial field generators (src/field_generator.h:11-24).
#pragma once
#include <cmath>

namespace field_generator {

// (300) lamellar W-(r) = A cos(2 pi (kx x/mx + ky y/my + kz z/mz)), w+ = 0, row-major index.
inline void create_lamellar(double* w, double Amplitude, const int* m, int kx = 3, int ky = 0, int kz = 0) {
    const long M = (long)m[0] * m[1] * m[2];
    for (int x = 0; x < m[0]; x++)
        for (int y = 0; y < m[1]; y++)
            for (int z = 0; z < m[2]; z++) {
                const long r = (long)m[2] * ((long)x * m[1] + y) + z;
                w[r] = Amplitude * std::cos(2.0 * M_PI * (kx * ((double)x) / m[0] + ky * ((double)y) / m[1] +
                                                          kz * ((double)z) / m[2]));
                w[r + M] = 0.0;
            }
}

}  // namespace field_generator
