/* psy_file.h — .psy scene files for the ray-tracing path.
 *
 * Reads the reference's scene description language into the descriptors of
 * psy_rt.h and psy_render.h, so an exported scene reaches psy_scene_create,
 * psy_renderer_create and psy_render unchanged:
 *
 *   psy_file_parse      DataTree::from_str (src/parse/data_tree.rs:22-45)
 *                       followed by parse_scene (src/parse/psy.rs:99-200) for
 *                       every `Scene` node of the file (src/main.rs:235-251)
 *   psy_file_get_scene  the parsed Renderer's scene, camera, world and
 *                       settings (src/renderer.rs:32-38, src/scene/mod.rs)
 *   psy_file_dump_tree  the DataTree of a text in a canonical line form, for
 *                       checking the tokenizer (no counterpart)
 *
 * Behaviour follows the reference reader: the same grammar (whitespace,
 * `#` comments, `$` identifiers and leaf contents with backslash escapes),
 * the same section rules, column-major 4x4 transforms (psy.rs:556-572), fan
 * triangulation of polygons (psy_mesh_surface.rs:96-115), colours `rec709`
 * (converted to XYZ with equal-energy white), `blackbody` and
 * `color_temperature` (psy.rs:577-615), and shader binding by instance.  The
 * reference's parse errors (PsyParseError, DataTree ParseError) and its
 * panics on malformed data (duplicate names, unknown shaders, short face
 * lists) come back as PSY_E_PARSE with psy_last_error() giving
 * "<kind> at byte <offset> (line <n>): <message>".
 *
 * Nested assemblies stay nested: psy_file_scene.scene is the root assembly
 * with its sub-assemblies in `assemblies`, each holding its own meshes,
 * lights and instances (instance indices are per assembly), as
 * AssemblyBuilder::add_assembly keeps them (scene/assembly.rs:197-282); the
 * world's shader list holds every assembly's shaders, the root's first, then
 * depth first in definition order.  A light instance of the root assembly
 * without a transform gets the identity.  Not supported (PSY_E_PARSE with a
 * message saying so): more than 64 time samples of a camera lens or light
 * parameter, and light
 * colour samples that mix representations (the reference panics when it
 * lerps them).  Every other time sample is kept.  Files carry no
 * bounce limit: settings.max_bounces is the reference renderer's fixed 2
 * (src/renderer.rs:599).
 */
#ifndef PSY_FILE_H
#define PSY_FILE_H

#include <stdint.h>

#include "psy_render.h"
#include "psy_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PSY_E_PARSE (-6) /* malformed scene file */

typedef struct psy_file psy_file;

/* Descriptors of one parsed scene; every pointer is owned by the psy_file
 * and valid until psy_file_destroy. */
typedef struct {
    const char* name;          /* the Scene's identifier without '$', or "" */
    const char* output_path;   /* Output { Path ["..."] } without quotes */
    psy_scene_desc scene;
    psy_world_desc world;
    psy_camera_desc camera;
    psy_render_settings settings;
} psy_file_scene;

/* Parses `len` bytes of UTF-8 text.  Fails on the first error. */
int psy_file_parse(const char* text, uint64_t len, psy_file** out);
int psy_file_destroy(psy_file* file);
/* Number of Scene nodes in the file. */
uint32_t psy_file_scene_count(const psy_file* file);
int psy_file_get_scene(const psy_file* file, uint32_t index, psy_file_scene* out);

/* Writes the DataTree of the text, one node per line: internal nodes as
 * "I <type> <ident or -> <byte offset>", leaves as "L <type> <byte offset>
 * <contents>" with backslash, newline and carriage return escaped as \\, \n
 * and \r, children indented by two spaces per level.  Writes at most `cap`
 * bytes to `out` (may be NULL when cap is 0) and returns the full length,
 * or a negative PSY_E* code. */
int64_t psy_file_dump_tree(const char* text, uint64_t len, char* out, uint64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* PSY_FILE_H */
