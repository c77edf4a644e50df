// TEST INFRASTRUCTURE: drives the library's host code (the .psy reader, the
// scene packer and the BVH4 builder) under AddressSanitizer and
// UndefinedBehaviorSanitizer.  Built and run by tests/test_sanitize.py; any
// out-of-bounds access, use after free, leak or undefined operation aborts
// the run with a report.
//
//   sanitize_main <edits-per-file> <file.psy>...
//
// Every file is parsed as is and after `edits` random edits of 1-4
// characters each (deletions, insertions and replacements from the grammar's
// alphabet, as the reader's fuzz test makes them); each scene a parse yields
// is packed (BVH4 builds, light tree, transforms).  Then seeded triangle
// soups, time-sampled meshes and instanced scenes go through the packer.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/psy_file.h"
#include "../../psychopath_amd/csrc/scene.h"

namespace psy {
namespace api {
thread_local std::string g_error;
int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}
}  // namespace api

// This build holds the library's host code only: the device builder is not
// linked, and the scenes built here never ask for it.
bool build_bvh4_device(const BuildInput&, uint32_t, HostBvh4&, std::string& why, double*) {
    why = "host-only build";
    return false;
}
bool build_light_bvh2_device(const BuildInput&, ArrayBvh2&, std::vector<uint32_t>&, std::string& why) {
    why = "host-only build";
    return false;
}
}  // namespace psy

namespace {

uint64_t g_state = 0x9e3779b97f4a7c15ull;
uint64_t next_u64() {
    // SplitMix64
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
uint64_t below(uint64_t n) { return n ? next_u64() % n : 0; }
float unit() { return (float)(next_u64() >> 40) / (float)(1ull << 24); }

int pack_file(const std::string& text, uint64_t& scenes_packed) {
    psy_file* f = nullptr;
    int rc = psy_file_parse(text.data(), text.size(), &f);
    std::vector<char> tree(1 << 16);
    (void)psy_file_dump_tree(text.data(), text.size(), tree.data(), tree.size());
    if (rc != PSY_OK) return rc;
    const uint32_t n = psy_file_scene_count(f);
    for (uint32_t i = 0; i < n; i++) {
        psy_file_scene sc{};
        if (psy_file_get_scene(f, i, &sc) != PSY_OK) continue;
        psy::PackedScene ps;
        std::string err;
        if (psy::pack_scene(&sc.scene, ps, err) == PSY_OK) scenes_packed++;
    }
    psy_file_destroy(f);
    return PSY_OK;
}

// A seeded soup of `tris` triangles with `samples` time samples, placed by
// `instances` instances with 0-3 transform samples each.
int pack_synthetic(uint32_t tris, uint32_t samples, uint32_t instances) {
    const uint32_t verts = 3 * tris;
    std::vector<float> v((size_t)samples * verts * 3);
    for (uint32_t t = 0; t < tris; t++) {
        const float cx = unit() * 20.0f - 10.0f, cy = unit() * 20.0f - 10.0f, cz = unit() * 20.0f - 10.0f;
        for (uint32_t s = 0; s < samples; s++)
            for (int k = 0; k < 3; k++) {
                float* p = &v[((size_t)s * verts + 3 * t + k) * 3];
                p[0] = cx + unit() + 0.1f * s;
                p[1] = cy + unit();
                p[2] = cz + unit();
                if (below(50) == 0) p[0] = cx;  // coincident coordinates
            }
    }
    std::vector<uint32_t> idx(verts);
    for (uint32_t i = 0; i < verts; i++) idx[i] = i;
    if (tris > 4 && below(3) == 0) idx[3] = idx[0];  // a degenerate triangle
    psy_mesh_desc m{verts, samples, v.data(), nullptr, tris, idx.data()};
    std::vector<psy_instance_desc> inst(instances);
    std::vector<float> xf;
    for (uint32_t i = 0; i < instances; i++) {
        const uint32_t k = (uint32_t)below(4);
        inst[i] = psy_instance_desc{0, PSY_NONE, (uint32_t)(xf.size() / 12), k, PSY_OBJECT_MESH};
        for (uint32_t j = 0; j < k; j++) {
            const float a = unit() * 6.0f;
            const float c = (float)std::cos(a), s = (float)std::sin(a);
            const float vals[12] = {c, -s, 0, unit() * 30, s, c, 0, unit() * 30, 0, 0, 1, unit() * 30};
            xf.insert(xf.end(), vals, vals + 12);
        }
    }
    psy_scene_desc d{};
    d.mesh_count = 1;
    d.meshes = &m;
    d.instance_count = instances;
    d.instances = inst.data();
    d.xform_count = (uint32_t)(xf.size() / 12);
    d.xforms = xf.empty() ? nullptr : xf.data();
    psy::PackedScene ps;
    std::string err;
    return psy::pack_scene(&d, ps, err);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <edits-per-file> <file.psy>...\n", argv[0]);
        return 2;
    }
    const int edits = std::atoi(argv[1]);
    const char alphabet[] = "{}[]$#\\ \n\tab0.-e\"";
    uint64_t parsed = 0, rejected = 0, packed = 0;
    for (int a = 2; a < argc; a++) {
        std::ifstream in(argv[a], std::ios::binary);
        std::stringstream ss;
        ss << in.rdbuf();
        const std::string base = ss.str();
        if (base.empty()) {
            std::fprintf(stderr, "cannot read %s\n", argv[a]);
            return 2;
        }
        for (int e = -1; e < edits; e++) {
            std::string t = base;
            if (e >= 0) {
                const int n = 1 + (int)below(4);
                for (int k = 0; k < n && !t.empty(); k++) {
                    const size_t i = below(t.size());
                    const uint64_t op = below(10);
                    if (op < 4) t.erase(i, 1);
                    else if (op < 8) t.insert(t.begin() + (long)i, alphabet[below(sizeof(alphabet) - 1)]);
                    else t[i] = alphabet[below(sizeof(alphabet) - 1)];
                }
            }
            if (pack_file(t, packed) == PSY_OK) parsed++;
            else rejected++;
        }
    }
    uint64_t synthetic = 0;
    const uint32_t sizes[] = {1, 2, 3, 4, 7, 50, 1000, 20000};
    for (uint32_t n : sizes)
        for (uint32_t s = 1; s <= 3; s++) {
            if (pack_synthetic(n, s, 1 + (uint32_t)below(40)) != PSY_OK) {
                std::fprintf(stderr, "synthetic scene %u x %u failed: %s\n", n, s, psy::api::g_error.c_str());
                return 1;
            }
            synthetic++;
        }
    std::printf("sanitize ok: %llu texts parsed, %llu rejected, %llu scenes packed, %llu synthetic scenes\n",
                (unsigned long long)parsed, (unsigned long long)rejected, (unsigned long long)packed,
                (unsigned long long)synthetic);
    return 0;
}
