// Host-side colour-space constants of the reference's colour crate
// (sub_crates/color/build.rs): the build script derives the RGB<->XYZ
// matrices in f64 from chromaticities (RGBtoXYZ of the ACES CTL,
// build.rs:178-225; Gauss-Jordan inverse of Imath's gjInverse,
// build.rs:229-300), prints them with ten decimals and the compiler reads
// them as f32 literals.  Rows are applied as (x*m0 + y*m1) + z*m2 in f32.
#pragma once

#include <cstdio>
#include <cstdlib>

namespace psy {
namespace color {

struct Matrices {
    float rec709_e_to_xyz[9];  // equal-energy white
    float xyz_to_rec709_e[9];
    float rec709_to_xyz[9];    // D65 white

    Matrices() {
        double m[3][3];
        rgb_to_xyz(m, 1.0 / 3.0, 1.0 / 3.0);
        double inv[3][3];
        gj_inverse(m, inv);
        double d65[3][3];
        rgb_to_xyz(d65, 0.3127, 0.3290);
        for (int i = 0; i < 9; i++) {
            rec709_e_to_xyz[i] = literal(m[i / 3][i % 3]);
            xyz_to_rec709_e[i] = literal(inv[i / 3][i % 3]);
            rec709_to_xyz[i] = literal(d65[i / 3][i % 3]);
        }
    }

private:
    static float literal(double v) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.10f", v);
        return std::strtof(buf, nullptr);
    }
    // Rec.709 primaries with white point (wx, wy).
    static void rgb_to_xyz(double mat[3][3], double wx, double wy) {
        const double rx = 0.640, ry = 0.330, gx = 0.300, gy = 0.600, bx = 0.150, by = 0.060;
        const double Y = 1.0;
        const double X = wx * Y / wy, Z = (1.0 - wx - wy) * Y / wy;
        const double d = rx * (by - gy) + bx * (gy - ry) + gx * (ry - by);
        const double sr = (X * (by - gy) - gx * (Y * (by - 1.0) + by * (X + Z)) + bx * (Y * (gy - 1.0) + gy * (X + Z))) / d;
        const double sg = (X * (ry - by) + rx * (Y * (by - 1.0) + by * (X + Z)) - bx * (Y * (ry - 1.0) + ry * (X + Z))) / d;
        const double sb = (X * (gy - ry) - rx * (Y * (gy - 1.0) + gy * (X + Z)) + gx * (Y * (ry - 1.0) + ry * (X + Z))) / d;
        mat[0][0] = sr * rx;
        mat[0][1] = sg * gx;
        mat[0][2] = sb * bx;
        mat[1][0] = sr * ry;
        mat[1][1] = sg * gy;
        mat[1][2] = sb * by;
        mat[2][0] = sr * (1.0 - rx - ry);
        mat[2][1] = sg * (1.0 - gx - gy);
        mat[2][2] = sb * (1.0 - bx - by);
    }
    // Gauss-Jordan elimination with partial pivoting on the first two
    // columns, then back substitution, in the build script's order.
    static void gj_inverse(const double m[3][3], double s[3][3]) {
        double t[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                t[i][j] = m[i][j];
                s[i][j] = i == j ? 1.0 : 0.0;
            }
        for (int i = 0; i < 2; i++) {
            int pivot = i;
            double size = t[i][i] < 0.0 ? -t[i][i] : t[i][i];
            for (int j = i + 1; j < 3; j++) {
                const double a = t[j][i] < 0.0 ? -t[j][i] : t[j][i];
                if (a > size) {
                    pivot = j;
                    size = a;
                }
            }
            if (pivot != i)
                for (int j = 0; j < 3; j++) {
                    double tmp = t[i][j];
                    t[i][j] = t[pivot][j];
                    t[pivot][j] = tmp;
                    tmp = s[i][j];
                    s[i][j] = s[pivot][j];
                    s[pivot][j] = tmp;
                }
            for (int j = i + 1; j < 3; j++) {
                const double f = t[j][i] / t[i][i];
                for (int k = 0; k < 3; k++) {
                    t[j][k] -= f * t[i][k];
                    s[j][k] -= f * s[i][k];
                }
            }
        }
        for (int i = 2; i >= 0; i--) {
            const double f = t[i][i];
            for (int j = 0; j < 3; j++) {
                t[i][j] /= f;
                s[i][j] /= f;
            }
            for (int j = 0; j < i; j++) {
                const double g = t[j][i];
                for (int k = 0; k < 3; k++) {
                    t[j][k] -= g * t[i][k];
                    s[j][k] -= g * s[i][k];
                }
            }
        }
    }
};

inline const Matrices& matrices() {
    static const Matrices m;
    return m;
}

inline void apply(const float m[9], float a, float b, float c, float out[3]) {
    out[0] = (a * m[0] + b * m[1]) + c * m[2];
    out[1] = (a * m[3] + b * m[4]) + c * m[5];
    out[2] = (a * m[6] + b * m[7]) + c * m[8];
}

}  // namespace color
}  // namespace psy
