/* psy_image.h — film output of the ray-tracing path.
 *
 * Encodes the XYZ film a render produces (psy_render, row 0 at the top)
 * the way the reference writes its Image:
 *
 *   PSY_IMAGE_PPM_ASCII  Image::write_ascii_ppm  (src/image.rs:101-119)
 *   PSY_IMAGE_PPM        Image::write_binary_ppm (src/image.rs:121-136)
 *   PSY_IMAGE_PNG        Image::write_png        (src/image.rs:138-164)
 *   PSY_IMAGE_EXR        Image::write_exr        (src/image.rs:166-195)
 *
 * PPM and PNG pixels are xyz_to_rec709_e, the sRGB transfer curve and
 * quantize_tri_255 (image.rs:282-312); PNG rows run from the film's last
 * row to its first, as the reference writes them, with alpha 255.  EXR
 * pixels are xyz_to_rec709_e as half floats in channels R, G, B; the file is
 * uncompressed where the reference uses PIZ (same pixels, larger file).
 *
 * psy_image_bucket_base64 is Bucket::rgba_base64(xyz_to_rec709_e)
 * (image.rs:227-252), the payload of the reference's Blender bucket
 * protocol (src/renderer.rs:312-346): four little-endian f32 per pixel
 * (R, G, B, 1.0) in scanline order, standard base64 with padding.
 *
 * Host code; `xyz` is a host pointer to width*height*3 floats.  Both
 * functions write at most `cap` bytes to `out` (which may be NULL when cap
 * is 0) and return the full size, or a negative PSY_E* code.
 */
#ifndef PSY_IMAGE_H
#define PSY_IMAGE_H

#include <stdint.h>

#include "psy_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PSY_IMAGE_PPM_ASCII 0u
#define PSY_IMAGE_PPM 1u
#define PSY_IMAGE_PNG 2u
#define PSY_IMAGE_EXR 3u

int64_t psy_image_encode(const float* xyz, uint32_t width, uint32_t height, uint32_t format, uint8_t* out,
                         uint64_t cap);

/* The bucket [x0, x1) x [y0, y1) of the film (clipped to the image). */
int64_t psy_image_bucket_base64(const float* xyz, uint32_t width, uint32_t height, uint32_t x0, uint32_t y0,
                                uint32_t x1, uint32_t y1, char* out, uint64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* PSY_IMAGE_H */
