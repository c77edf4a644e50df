"""Command line of the reference renderer (src/main.rs:62-355) over the GPU
path: `python -m psychopath_amd -i scene.psy [-s N] [--crop X1 Y1 X2 Y2]
[--stats] [--serialized_output] [--use_stdin] [--dev]`.

Every Scene of the file is built, rendered on the GPU and written to its
Output path (.png or .exr, as Image::write_png / write_exr).  `-b/--spb` and
`-t/--threads` are accepted for compatibility: the GPU renders in waves of
paths instead of CPU buckets and threads.  `--crop` renders only the
inclusive rectangle and leaves the rest black, as the reference does.
With `--serialized_output` the image goes to standard output in the
reference's bucket protocol (DIV / percentage / rectangle / base64 RGBA /
BUCKET_END / DIV, renderer.rs:328-336) instead of to disk.
"""

from __future__ import annotations

import argparse
import sys
import time


def _read_stdin() -> str:
    """The scene text up to the __PSY_EOF__ marker (main.rs:194-225)."""
    data = b""
    while True:
        chunk = sys.stdin.buffer.read1(4096) if hasattr(sys.stdin.buffer, "read1") else sys.stdin.buffer.read(4096)
        if not chunk:
            raise SystemExit("Unexpected end of scene input.")
        data += chunk
        at = data.find(b"__PSY_EOF__")
        if at >= 0:
            return data[:at].decode("utf-8")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m psychopath_amd", description="A slightly psychotic path tracer "
                                 "(MI355X ray-tracing path)")
    ap.add_argument("-i", "--input", metavar="FILE", help="Input .psy file")
    ap.add_argument("-s", "--spp", type=int, metavar="N", help="Number of samples per pixel")
    ap.add_argument("-b", "--spb", type=int, metavar="N", help="Target number of samples per bucket (accepted; the "
                    "GPU renders in waves of paths)")
    ap.add_argument("--crop", type=int, nargs=4, metavar=("X1", "Y1", "X2", "Y2"),
                    help="Only keep the image between pixel coordinates (X1, Y1) and (X2, Y2), inclusive")
    ap.add_argument("-t", "--threads", type=int, metavar="N", help="Accepted; the render runs on the GPU")
    ap.add_argument("--stats", action="store_true", help="Print additional statistics about rendering")
    ap.add_argument("--dev", action="store_true", help="Show useful dev/debug info")
    ap.add_argument("--serialized_output", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--use_stdin", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args(argv)

    from . import film, psyfile, render

    if a.dev:
        # Record sizes of the GPU layout (the reference prints its structs').
        print("Ray record (SoA):  33 bytes (origin, direction, time, max_t, flags)")
        print("Hit record (SoA):  25 bytes (kind, t, 3 barycentrics, prim id, instance id)")
        print("BVH4 node record:  128 bytes (4 child boxes, refs, visit orders, meta)")
        print("Triangle record:   48 bytes static, 36 per time sample + 4 with motion")
        return 0
    if not a.input and not a.use_stdin:
        ap.error("the following arguments are required: -i/--input")
    if a.crop and (a.crop[0] > a.crop[2] or a.crop[1] > a.crop[3]):
        ap.error("Argument '--crop': X1 must be less than or equal to X2 and Y1 less than or equal to Y2")
    say = (lambda *x: None) if a.serialized_output else (lambda *x: print(*x, flush=True))

    say("Parsing scene file...")
    t = time.perf_counter()
    text = _read_stdin() if a.use_stdin else open(a.input, encoding="utf-8").read()
    scenes = psyfile.read(text)
    say(f"\tParsed scene file in {time.perf_counter() - t:.3f}s")
    for sc in scenes:
        say("Building scene...")
        t = time.perf_counter()
        if a.spp is not None:
            say(f"\tOverriding scene spp: {a.spp}")
            sc.settings["spp"] = a.spp
        r = render.build(sc)
        say(f"\tBuilt scene in {time.perf_counter() - t:.3f}s")
        settings = dict(sc.settings)
        if a.crop:
            # Inclusive corners clamped to the image (renderer.rs:97-108).
            w, h = settings["width"], settings["height"]
            x1, y1, x2, y2 = (min(a.crop[0], w - 1), min(a.crop[1], h - 1), min(a.crop[2], w - 1),
                              min(a.crop[3], h - 1))
            settings["crop"] = (x1, y1, x2 - x1 + 1, y2 - y1 + 1)
        say("Rendering scene on the GPU...")
        t = time.perf_counter()
        img, st = r.render(sc.camera, settings)
        rtime = time.perf_counter() - t
        say(f"\tRendered scene in {rtime:.3f}s")
        say(f"\t\tTrace:                  {st['trace_ms'] / 1e3:.3f}s (device time of the whole render)")
        say(f"\t\t\tRays traced:          {st['rays']}")
        say(f"\t\t\tRays/sec:             {int(st['rays'] / max(rtime, 1e-9))}")
        if a.stats:
            say(f"\t\tPaths:                  {st['paths']} in {st['waves']} waves")
        if a.serialized_output:
            for block in film.blender_buckets(img):
                sys.stdout.write(block)
            sys.stdout.flush()
            continue
        say(f"Writing image to disk into '{sc.output_path}'...")
        t = time.perf_counter()
        if not (sc.output_path.endswith(".png") or sc.output_path.endswith(".exr")):
            raise SystemExit("Unknown output file extension.")
        film.write(sc.output_path, img)
        say(f"\tWrote image in {time.perf_counter() - t:.3f}s")
        del r
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
