"""Command-line renderer: the headless counterpart of the reference's main()
(src/main.cpp:10-32) and PathTracer::render(num) (src/renderer/pathtracer.cu:152-170).

    python -m gpr SCENE [--spp N] [--depth D] [--rng reference|pcg] [--seed S]
                        [--rr-start R] [--traversal fast|ref] [--out FILE.png|FILE.hdr]
                        [--resume CKPT.npz] [--checkpoint CKPT.npz]

Loads a Mitsuba-XML / .obj / .gprs scene, renders N samples (progress on stderr,
as the reference prints it), reports the time spent and writes the mean image
(default "<samples>spp.png", the reference's file name).  --checkpoint saves the
linear radiance sums and the sample count; --resume continues from such a file
(PCG mode), so a render can be split across runs with an identical result.
"""
import argparse
import sys
import time

import numpy as np

from . import (GprError, RNG_PCG, RNG_REFERENCE, Renderer, Scene, TRAVERSAL_FAST, TRAVERSAL_REF,
               synchronize)


def parse(argv):
    ap = argparse.ArgumentParser(prog="python -m gpr", description=__doc__.split("\n\n")[0])
    ap.add_argument("scene", help="scene file (.xml, .obj or .gprs)")
    ap.add_argument("--spp", type=int, default=100, help="samples per pixel to render (default 100)")
    ap.add_argument("--depth", type=int, default=33, help="maximum path length (PTOptions default 33)")
    ap.add_argument("--rng", choices=["reference", "pcg"], default="reference",
                    help="reference: TEA/LCG as the reference; pcg: per-path PCG32 streams")
    ap.add_argument("--seed", type=int, default=1, help="PCG seed")
    ap.add_argument("--rr-start", type=int, default=3, help="first bounce of Russian roulette (PCG mode)")
    ap.add_argument("--traversal", choices=["fast", "ref"], default="fast")
    ap.add_argument("--out", default=None, help="output image, .png or .hdr (default <samples>spp.png)")
    ap.add_argument("--resume", default=None, help="continue from a checkpoint written by --checkpoint")
    ap.add_argument("--checkpoint", default=None, help="save the radiance sums and sample count (.npz)")
    a = ap.parse_args(argv)
    if a.spp < 0 or a.depth < 1:
        ap.error("--spp must be >= 0 and --depth >= 1")
    return a


def main(argv=None):
    a = parse(sys.argv[1:] if argv is None else argv)
    t0 = time.perf_counter()
    try:
        scene = Scene.load(a.scene)
        r = Renderer(scene, max_depth=a.depth, rng_mode=RNG_PCG if a.rng == "pcg" else RNG_REFERENCE,
                     rr_start=a.rr_start, seed=a.seed,
                     traversal=TRAVERSAL_FAST if a.traversal == "fast" else TRAVERSAL_REF)
        done = 0
        if a.resume:
            ck = np.load(a.resume, allow_pickle=False)
            r.restore(ck["rgb_sum"], int(ck["spp"]))
            done = int(ck["spp"])
        for i in range(a.spp):
            r.render(1)
            sys.stderr.write("\r%f%%" % (100.0 * (i + 1) / a.spp))
            sys.stderr.flush()
        synchronize()
        if a.spp:
            sys.stderr.write("\n")
        total = done + a.spp
        print("SpendTime = %.3fs" % (time.perf_counter() - t0), flush=True)
        out = a.out or ("%dspp.png" % total)
        r.write_image(out)
        if a.checkpoint:
            rgb, _ = r.read()
            np.savez(a.checkpoint, rgb_sum=rgb, spp=np.uint32(total))
        print("wrote %s (%d samples)" % (out, total), flush=True)
        r.free()
        scene.free()
    except GprError as e:
        print("error: %s" % e, file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
