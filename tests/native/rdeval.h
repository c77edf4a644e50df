// TEST INFRASTRUCTURE: one evaluation of a renderer math function from a
// flat float input record, shared by the host (rdcheck.cpp) and device
// (rdgpu.hip) builds so both compute exactly the same calls.
//   which  inputs (IN_STRIDE floats)                    outputs (OUT_STRIDE floats)
//   0  closure sample: kind rough fres rgb inc nor nor_g u v wl   dir filter pdf
//   1  closure evaluate: kind rough fres rgb inc out nor nor_g wl  filter pdf
//   2  sphere estimate: kind rough fres rgb inc to r2 nor nor_g    value
//      (closures: the dielectric's ior at [28], its sample number w at [29])
//   3  rect light sample: w h rgb xf[12] arr u v wl               color point normal err pdf, sample_pdf
//   4  sphere light sample: r temp xf[12] arr u v wl              color point normal err pdf, sample_pdf
//   5  colour: kind a b c wl                                      spectrum xyz
//   6  spherical triangle: a b c u v                              area sample
//   7  concentric disc + cosine hemisphere + sphere: u v           disc(2) hemi(3) sphere(3)
//   8  probit: n width                                            value
#pragma once

#define RDEVAL_IN_STRIDE 32
#define RDEVAL_OUT_STRIDE 16

namespace rdeval {

using namespace psy;
using namespace psy::rd;

RD V3 in3(const float* p) { return v3(p[0], p[1], p[2]); }
RD void out3(float* p, V3 a) {
    p[0] = a.x;
    p[1] = a.y;
    p[2] = a.z;
}
RD void out4(float* p, S4 a) {
    for (int k = 0; k < 4; k++) p[k] = a.v[k];
}
RD GpuClosure clo(const float* p) {
    GpuClosure c;
    c.kind = (uint32_t)p[0];
    c.roughness = p[1];
    c.fresnel = p[2];
    c.ior = p[28];
    c.color = GpuColor{PSY_COLOR_XYZ, p[3], p[4], p[5]};
    return c;
}

RD void eval(int which, const float* in, float* o) {
    for (int k = 0; k < RDEVAL_OUT_STRIDE; k++) o[k] = 0.0f;
    if (which == 0) {
        V3 d;
        S4 f;
        float pdf;
        closure_sample(clo(in), in3(in + 6), in3(in + 9), in3(in + 12), in[15], in[16], in[29], in[17], d, f, pdf);
        out3(o, d);
        out4(o + 3, f);
        o[7] = pdf;
    } else if (which == 1) {
        S4 f;
        float pdf;
        closure_evaluate(clo(in), in3(in + 6), in3(in + 9), in3(in + 12), in3(in + 15), in[18], f, pdf);
        out4(o, f);
        o[4] = pdf;
    } else if (which == 2) {
        o[0] = closure_estimate_sphere_light(clo(in), in3(in + 6), in3(in + 9), in[12], in3(in + 13), in3(in + 16));
    } else if (which == 3) {
        GpuRectLight L{in[0], in[1], 0u, 0u, GpuColor{PSY_COLOR_XYZ, in[2], in[3], in[4]}};
        const Xf x = xf_from_values(in + 5);
        const LightSample s = rect_sample_from_point(L, x, in3(in + 17), in[20], in[21], in[22]);
        out4(o, s.color);
        out3(o + 4, s.point);
        out3(o + 7, s.normal);
        o[10] = s.point_err;
        o[11] = s.pdf;
        o[12] = rect_sample_pdf(L, x, in3(in + 17), in3(in + 23), in3(in + 26));
    } else if (which == 4) {
        GpuSphereLight L{in[0], 0u, 0u, 0.0f, GpuColor{PSY_COLOR_TEMPERATURE, in[1], 40.0f, 0.0f}};
        const Xf x = xf_from_values(in + 2);
        const LightSample s = sphere_sample_from_point(L, x, in3(in + 14), in[17], in[18], in[19]);
        out4(o, s.color);
        out3(o + 4, s.point);
        out3(o + 7, s.normal);
        o[10] = s.point_err;
        o[11] = s.pdf;
        o[12] = sphere_sample_pdf(L, x, in3(in + 14));
    } else if (which == 5) {
        const GpuColor c{(uint32_t)in[0], in[1], in[2], in[3]};
        const S4 s = color_spectral(c, in[4]);
        out4(o, s);
        out3(o + 4, xyz_from_spectral(s, in[4]));
    } else if (which == 6) {
        o[0] = spherical_triangle_solid_angle(in3(in), in3(in + 3), in3(in + 6));
        out3(o + 1, uniform_sample_spherical_triangle(in3(in), in3(in + 3), in3(in + 6), in[9], in[10]));
    } else if (which == 7) {
        float dx, dy;
        square_to_circle(in[0] * 2.0f - 1.0f, in[1] * 2.0f - 1.0f, dx, dy);
        o[0] = dx;
        o[1] = dy;
        out3(o + 2, cosine_sample_hemisphere(in[0], in[1]));
        out3(o + 5, uniform_sample_sphere(in[0], in[1]));
    } else {
        o[0] = probit(in[0], in[1]);
    }
}

}  // namespace rdeval
