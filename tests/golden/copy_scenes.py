"""Copies a set of the reference's test scene files (data: NFF scenes and the OBJ / COLLADA meshes
they name, from testing/scenes of the reference) into tests/golden/scenes/, so the scene-loader
tests run without the reference.  Run from the repository root where the reference is mounted:
    python tests/golden/copy_scenes.py /root/reference
"""
import os
import shutil
import sys

SCENES = [
    "cornelbox_translucentsphere.nff", "cornelbox.dae", "unitrectangle.obj", "unitrectangle.nff", "medium_in_medium.nff",
    "unitcube.dae", "veach_mis_scene_distant.nff", "veach_mis_scene_distant.obj", "nested_media1.nff", "litcorner.nff",
    "env_sphere_mirror.nff", "unitdisc_24_xz.obj", "cubelight.nff", "off_center_rgb_cmy.nff", "medium_backlighting_area.nff",
    "atmosphere.nff", "earth_atmosphere_collision_coefficients.json",
]

if __name__ == "__main__":
    src = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "testing", "scenes")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")
    os.makedirs(dst, exist_ok=True)
    for f in SCENES:
        shutil.copy(os.path.join(src, f), os.path.join(dst, f))
        print(f)
