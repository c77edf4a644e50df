// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Lights: restates src/light/rectangle_light.rs, src/light/sphere_light.rs
// and src/light/distant_disk_light.rs (one time sample of size and colour
// each).
#pragma once

#include "psyref_color.h"
#include "psyref_math.h"
#include "psyref_sample.h"
#include "psyref_bbox.h"
#include "psyref_tri.h"

namespace psyref {

struct LightSample {
    Vec4 color;
    Vec3 point, normal;
    float point_err;
    float pdf;
};

// ---- rectangle_light.rs --------------------------------------------------
struct RectangleLight {
    float w = 1, h = 1;
    Color color;
    // Time samples (rectangle_light.rs:24-29): (w, h) pairs and colours,
    // each list lerped on its own at the ray's time; empty = the one value
    // above.
    std::vector<float> dims_t;
    std::vector<Color> colors_t;
    static const float SIMPLE_SAMPLING_THRESHOLD;

    // The light at `time`, with one value per field.
    RectangleLight at(float time) const {
        RectangleLight r;
        r.w = w;
        r.h = h;
        r.color = color;
        const size_t nd = dims_t.size() / 2;
        if (nd > 0 && !(nd == 1 || time == 1.0f)) {
            const float tmp = time * (float)(nd - 1);
            const size_t i1 = (size_t)tmp;
            const float a2 = tmp - (float)i1;
            r.w = lerp(dims_t[2 * i1], dims_t[2 * i1 + 2], a2);
            r.h = lerp(dims_t[2 * i1 + 1], dims_t[2 * i1 + 3], a2);
        } else if (nd > 0) {
            r.w = dims_t[2 * nd - 2];
            r.h = dims_t[2 * nd - 1];
        }
        if (!colors_t.empty()) r.color = lerp_slice(colors_t.data(), colors_t.size(), time);
        return r;
    }
    BBox bounds() const { return BBox(Vec3(w * -0.5f, h * -0.5f, 0.0f), Vec3(w * 0.5f, h * 0.5f, 0.0f)); }
    // One box per dimensions sample (rectangle_light.rs:36-42).
    std::vector<BBox> bounds_list() const {
        if (dims_t.empty()) return {bounds()};
        std::vector<BBox> out;
        for (size_t i = 0; i + 1 < dims_t.size(); i += 2)
            out.push_back(BBox(Vec3(dims_t[i] * -0.5f, dims_t[i + 1] * -0.5f, 0.0f),
                               Vec3(dims_t[i] * 0.5f, dims_t[i + 1] * 0.5f, 0.0f)));
        return out;
    }
    float approximate_energy() const { return colors_t.empty() ? mean_energy({color}) : mean_energy(colors_t); }

    void corners(const Transform& space_inv, Vec3 p[4]) const {
        p[0] = xform_point(Vec3(w * 0.5f, h * 0.5f, 0.0f), space_inv);
        p[1] = xform_point(Vec3(w * -0.5f, h * 0.5f, 0.0f), space_inv);
        p[2] = xform_point(Vec3(w * -0.5f, h * -0.5f, 0.0f), space_inv);
        p[3] = xform_point(Vec3(w * 0.5f, h * -0.5f, 0.0f), space_inv);
    }

    // rectangle_light.rs:51-96
    float sample_pdf(const Transform& space, Vec3 arr, Vec3 sample_dir, Vec3 hit_point) const {
        Transform space_inv = xform_inverse(space);
        Vec3 p[4];
        corners(space_inv, p);
        Vec3 sp1 = normalized(p[0] - arr), sp2 = normalized(p[1] - arr), sp3 = normalized(p[2] - arr),
             sp4 = normalized(p[3] - arr);
        float area_1 = spherical_triangle_solid_angle(sp2, sp1, sp3);
        float area_2 = spherical_triangle_solid_angle(sp4, sp1, sp3);
        Vec3 normal = xform_normal(Vec3(0.0f, 0.0f, 1.0f), space_inv);
        if ((area_1 + area_2) < SIMPLE_SAMPLING_THRESHOLD) {
            float area = triangle_surface_area(p[1], p[0], p[2]) + triangle_surface_area(p[3], p[0], p[2]);
            return (length2(hit_point - arr) / std::fabs(dot(normalized(sample_dir), normalized(normal)))) / area;
        }
        return 1.0f / (area_1 + area_2);
    }

    // rectangle_light.rs:120-243
    LightSample sample_from_point(const Transform& space, Vec3 arr, float u, float v, float wavelength) const {
        double surface_area = (double)w * (double)h;
        double surface_area_inv = 1.0 / surface_area;
        Transform space_inv = xform_inverse(space);
        Vec3 p[4];
        corners(space_inv, p);
        Vec3 sp1 = normalized(p[0] - arr), sp2 = normalized(p[1] - arr), sp3 = normalized(p[2] - arr),
             sp4 = normalized(p[3] - arr);
        float area_1 = spherical_triangle_solid_angle(sp2, sp1, sp3);
        float area_2 = spherical_triangle_solid_angle(sp4, sp1, sp3);
        Vec3 normal = xform_normal(Vec3(0.0f, 0.0f, 1.0f), space_inv);
        LightSample s;
        s.point_err = 0.0001f;
        s.normal = normal;
        if ((area_1 + area_2) < SIMPLE_SAMPLING_THRESHOLD) {
            float sa1 = triangle_surface_area(p[1], p[0], p[2]);
            float sa2 = triangle_surface_area(p[3], p[0], p[2]);
            float thr = sa1 / (sa1 + sa2);
            Vec3 sp = u < thr ? uniform_sample_triangle(p[1], p[0], p[2], v, u / thr)
                              : uniform_sample_triangle(p[3], p[0], p[2], v, (u - thr) / (1.0f - thr));
            Vec3 shadow_vec = sp - arr;
            s.color = (color.to_spectral_sample(wavelength) * (float)surface_area_inv) * 0.5f;
            s.pdf = (length2(sp - arr) / std::fabs(dot(normalized(shadow_vec), normalized(normal)))) / (sa1 + sa2);
            s.point = sp;
            return s;
        }
        float prob_1 = std::isinf(area_1) ? 1.0f : (std::isinf(area_2) ? 0.0f : area_1 / (area_1 + area_2));
        float prob_2 = 1.0f - prob_1;
        Vec3 shadow_vec = u < prob_1 ? uniform_sample_spherical_triangle(sp2, sp1, sp3, v, u / prob_1)
                                     : uniform_sample_spherical_triangle(sp4, sp1, sp3, v,
                                                                         1.0f - ((u - prob_1) / prob_2));
        Vec3 arr_local = xform_point(arr, space);
        Vec3 sv_local = xform_vector(shadow_vec, space);
        sv_local = sv_local * (-arr_local.z / sv_local.z);
        Vec3 spl = arr_local + sv_local;
        // f32::max / f32::min (NaN-ignoring)
        auto fmaxn = [](float a, float b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); };
        auto fminn = [](float a, float b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); };
        spl.x = fminn(fmaxn(spl.x, w * -0.5f), w * 0.5f);
        spl.y = fminn(fmaxn(spl.y, h * -0.5f), h * 0.5f);
        spl.z = 0.0f;
        s.point = xform_point(spl, space_inv);
        s.pdf = 1.0f / (area_1 + area_2);
        s.color = (color.to_spectral_sample(wavelength) * (float)surface_area_inv) * 0.5f;
        return s;
    }

    // rectangle_light.rs:257-338 for one ray; returns true on a hit and fills
    // t, pos, pos_err, normal, sample_pdf.
    bool intersect(const Transform& xform, Vec3 orig, Vec3 dir, float max_t, float& t_out, Vec3& pos, float& pos_err,
                   Vec3& normal, float& sample_pdf_out) const {
        Transform space_inv = xform_inverse(xform);
        Vec3 p[4];
        corners(space_inv, p);
        RayTriPrecompute pre(dir);
        const Vec3 tris[2][3] = {{p[0], p[1], p[2]}, {p[2], p[3], p[0]}};
        for (int k = 0; k < 2; k++) {
            TriHit h;
            if (intersect_triangle(orig, pre, max_t, tris[k][0], tris[k][1], tris[k][2], h)) {
                if (h.t < max_t) {
                    surface_point(tris[k][0], tris[k][1], tris[k][2], h.b0, h.b1, h.b2, pos, pos_err);
                    normal = cross(tris[k][0] - tris[k][1], tris[k][0] - tris[k][2]);
                    sample_pdf_out = sample_pdf(xform, orig, dir, pos);
                    t_out = h.t;
                    return true;
                }
            }
        }
        return false;
    }
    Color emit_color() const {
        float inv_surface_area = (float)(1.0 / ((double)w * (double)h));
        return color.scaled(inv_surface_area);
    }
};

// ---- sphere_light.rs -------------------------------------------------------
struct SphereLight {
    float radius = 1;
    Color color;
    // Time samples (sphere_light.rs:25-31); empty = the one value above.
    std::vector<float> radii_t;
    std::vector<Color> colors_t;

    SphereLight at(float time) const {
        SphereLight r;
        r.radius = radii_t.empty() ? radius : lerp_slice(radii_t.data(), radii_t.size(), time);
        r.color = colors_t.empty() ? color : lerp_slice(colors_t.data(), colors_t.size(), time);
        return r;
    }
    BBox bounds() const { return BBox(Vec3(-radius, -radius, -radius), Vec3(radius, radius, radius)); }
    // One box per radius sample (sphere_light.rs:35-40).
    std::vector<BBox> bounds_list() const {
        if (radii_t.empty()) return {bounds()};
        std::vector<BBox> out;
        for (float r : radii_t) out.push_back(BBox(Vec3(-r, -r, -r), Vec3(r, r, r)));
        return out;
    }
    float approximate_energy() const { return colors_t.empty() ? mean_energy({color}) : mean_energy(colors_t); }

    // sphere_light.rs:51-81 (arr in world space)
    float sample_pdf(const Transform& space, Vec3 arr_world) const {
        Vec3 arr = xform_point(arr_world, space);
        double r = (double)radius;
        double d2 = (double)length2(Vec3(0, 0, 0) - arr);
        double d = std::sqrt(d2);
        if (d > r) {
            double s2 = (r * r) / d2;
            double sin_theta_max2 = s2 < 1.0 ? s2 : 1.0;
            double cos_theta_max = std::sqrt(1.0 - sin_theta_max2);
            return (float)uniform_sample_cone_pdf(cos_theta_max);
        }
        return (float)(1.0 / (4.0 * PI_64));
    }

    // sphere_light.rs:85-192
    LightSample sample_from_point(const Transform& space, Vec3 arr_world, float u, float v, float wavelength) const {
        Vec3 arr = xform_point(arr_world, space);
        Vec3 pos(0.0f, 0.0f, 0.0f);
        Transform inv_space = xform_inverse(space);
        double r = (double)radius;
        double surface_area_inv = 1.0 / (((4.0 * PI_64) * r) * r);
        Vec3 zv = pos - arr;
        double d2 = (double)length2(zv);
        double d = std::sqrt(d2);
        Vec3 z, x, y;
        coordinate_system_from_vector(zv, z, x, y);
        x = normalized(x);
        y = normalized(y);
        z = normalized(z);
        float sample_point_err = length(xform_vector(Vec3((float)r, (float)r, (float)r), inv_space)) * 0.001f;
        LightSample s;
        s.point_err = sample_point_err;
        if (d > r) {
            double s2 = (r * r) / d2;
            double sin_theta_max2 = s2 < 1.0 ? s2 : 1.0;
            double cos_theta_max2 = 1.0 - sin_theta_max2;
            double sin_theta_max = std::sqrt(sin_theta_max2);
            double cos_theta_max = std::sqrt(cos_theta_max2);
            Vec3 smp = normalized(uniform_sample_cone(u, v, cos_theta_max));
            double cos_theta = (double)smp.z;
            double cos_theta2 = cos_theta * cos_theta;
            double st2 = 1.0 - cos_theta2;
            double sin_theta2 = st2 > 0.0 ? st2 : 0.0;
            double sin_theta = std::sqrt(sin_theta2);
            double dd = 1.0 - (((d2 * sin_theta) * sin_theta) / (r * r));
            double cos_a = dd <= 0.0 ? sin_theta_max : (((d / r) * sin_theta2) + (cos_theta * std::sqrt(dd)));
            double sa2 = 1.0 - (cos_a * cos_a);
            double sin_a = std::sqrt(sa2 > 0.0 ? sa2 : 0.0);
            double phi = ((double)v * 2.0) * PI_64;
            Vec3 sv((float)((std::cos(phi) * sin_a) * r), (float)((std::sin(phi) * sin_a) * r),
                    (float)(d - (cos_a * r)));
            Vec3 sample_vec = ((x * sv.x) + (y * sv.y)) + (z * sv.z);
            Vec3 normal = normalized(arr + sample_vec);
            Vec3 point = normal * (float)r;
            s.point = xform_point(point, inv_space);
            s.normal = xform_normal(normal, inv_space);
            s.pdf = (float)uniform_sample_cone_pdf(cos_theta_max);
            s.color = color.to_spectral_sample(wavelength) * (float)surface_area_inv;
            return s;
        }
        Vec3 sample_vec = uniform_sample_sphere(u, v);
        Vec3 normal = normalized(arr + sample_vec);
        Vec3 point = normal * (float)r;
        s.point = xform_point(point, inv_space);
        s.normal = xform_normal(normal, inv_space);
        s.pdf = (float)(1.0 / (4.0 * PI_64));
        s.color = color.to_spectral_sample(wavelength) * (float)surface_area_inv;
        return s;
    }

    // sphere_light.rs:206-339 for one ray (orig_local = ray's local origin).
    bool intersect(const Transform& xform, Vec3 orig_world, Vec3 orig_local, Vec3 dir_world, float max_t,
                   float& t_out, Vec3& pos, float& pos_err, Vec3& normal, float& sample_pdf_out) const {
        Vec3 orig = orig_local;
        Vec3 dir = xform_vector(dir_world, xform);
        float a = length2(dir);
        float b = 2.0f * dot(dir, orig);
        float c = length2(orig) - (radius * radius);
        float disc = (b * b) - ((4.0f * a) * c);
        if (disc < 0.0f) return false;
        disc = std::sqrt(disc);
        float q = b < 0.0f ? -0.5f * (b - disc) : -0.5f * (b + disc);
        float t0 = q / a;
        float t1 = q != 0.0f ? c / q : max_t;
        if (t0 > t1) std::swap(t0, t1);
        if (t0 > max_t || t1 <= 0.0f) return false;
        float t;
        if (t0 > 0.0f) t = t0;
        else if (t1 <= max_t) t = t1;
        else return false;
        Transform inv_xform = xform_inverse(xform);
        Vec3 t_pos = orig + (dir * t);
        Vec3 unit_pos = normalized(t_pos);
        // `unit_pos * radius * inv_xform` is a Vector times a Transform in the
        // reference (transform_vector3a: no translation), kept as is.
        pos = xform_vector(unit_pos * radius, inv_xform);
        pos_err = 0.001f;
        normal = xform_normal(unit_pos, inv_xform);
        sample_pdf_out = sample_pdf(xform, orig_world);
        t_out = t;
        return true;
    }
    Color emit_color() const {
        float inv_surface_area = (float)(1.0 / (((4.0 * PI_64) * (double)radius) * (double)radius));
        return color.scaled(inv_surface_area);
    }
};

// ---- distant_disk_light.rs -------------------------------------------------
struct DistantDiskLight {
    float radius = 0.1f;
    Vec3 direction{0, 0, -1};
    Color color;
    // Time samples (distant_disk_light.rs:17-22); empty = the one value.
    std::vector<float> radii_t;
    std::vector<Vec3> directions_t;
    std::vector<Color> colors_t;

    DistantDiskLight at(float time) const {
        DistantDiskLight r;
        r.radius = radii_t.empty() ? radius : lerp_slice(radii_t.data(), radii_t.size(), time);
        r.direction = directions_t.empty() ? direction : lerp_slice(directions_t.data(), directions_t.size(), time);
        r.color = colors_t.empty() ? color : lerp_slice(colors_t.data(), colors_t.size(), time);
        return r;
    }
    float approximate_energy() const { return colors_t.empty() ? mean_energy({color}) : mean_energy(colors_t); }
    // distant_disk_light.rs:59-85: (color, shadow direction, pdf)
    void sample_from_point(float u, float v, float wavelength, Vec4& col, Vec3& shadow_vec, float& pdf) const {
        double r = (double)radius;
        double solid_angle_inv = 1.0 / ((2.0 * PI_64) * (1.0 - std::cos(r)));
        Vec3 z, x, y;
        coordinate_system_from_vector(neg(normalized(direction)), z, x, y);
        double cos_theta_max = std::cos(r);
        Vec3 smp = normalized(uniform_sample_cone(u, v, cos_theta_max));
        col = color.to_spectral_sample(wavelength) * (float)solid_angle_inv;
        shadow_vec = ((x * smp.x) + (y * smp.y)) + (z * smp.z);
        pdf = (float)uniform_sample_cone_pdf(cos_theta_max);
    }
};

}  // namespace psyref
