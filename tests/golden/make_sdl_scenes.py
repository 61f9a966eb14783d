"""Reads five of the reference's sample scenes (distribution/scenes:
bsp/Tango.pov, objects/mesh2.pov, radiosity/cornell.pov, objects/sor1.pov,
objects/lathe2.pov; include files from distribution/include) with
povray_amd.sdl and stores the parsed scenes as .npz fixtures beside this
script, for the GPU tests (the GPU machines have no copy of the reference).
The fixtures hold only the parsed geometry, camera, lights and plain
materials, not the scene files' text.  Run from the repository root:

    python tests/golden/make_sdl_scenes.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from povray_amd import sdl  # noqa: E402

SCENES = {"tango": ("bsp/Tango.pov", 200, 150), "mesh2": ("objects/mesh2.pov", 200, 150),
          "cornell": ("radiosity/cornell.pov", 160, 160), "sor1": ("objects/sor1.pov", 240, 180),
          "lathe2": ("objects/lathe2.pov", 240, 180)}
SRC = "/root/reference/distribution/scenes"
LIB = "/root/reference/distribution/include"

if __name__ == "__main__":
    for name, (rel, w, h) in SCENES.items():
        s = sdl.parse_file(os.path.join(SRC, rel), w, h, [LIB])
        s.name = f"sdl_{name}"
        s.save(os.path.join(HERE, f"sdl_{name}.npz"))
        print(name, s.n_objects, "objects;", len(s.settings.get("warnings", [])), "reductions")
