"""Writes the scene-file fixtures under tests/golden/scenes/ (run from the repo root:
`python tests/golden/make_scene_files.py`). Files follow the reference's JSON scene
format (models / instances / materials / camera / renderer / sky) with OBJ+MTL models:

* room.obj / room.mtl: an open box [0,1]^3 (floor, ceiling, back, left, right walls as
  quads) with a small ceiling light quad; per-face materials by name.
* block.obj: a unit cube centred at the origin, written with shared vertices, quads,
  negative (relative) indices and no normals (smooth normals are generated), one
  material defined inline-free (no mtllib: its faces use the default material).
* octa.obj: an octahedron with explicit vertex normals (vn) on some faces.
* box_scene.json: two blocks and an octahedron placed by transform operations, material
  overrides, camera, renderer and a disabled sky.
"""
import json
import os

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")


def room():
    v = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1),
         (0.375, 0.999, 0.375), (0.625, 0.999, 0.375), (0.625, 0.999, 0.625), (0.375, 0.999, 0.625)]
    lines = ["# open box, normals point inside", "mtllib room.mtl"]
    lines += ["v %r %r %r" % p for p in v]
    lines += ["usemtl white", "f 1 4 3 2"]            # floor (normal +y)
    lines += ["f 5 6 7 8"]                            # ceiling (normal -y)
    lines += ["f 4 8 7 3"]                            # back wall z=1 (normal -z)
    lines += ["usemtl red", "f 1 5 8 4"]              # left wall x=0 (normal +x)
    lines += ["usemtl green", "f 2 3 7 6"]            # right wall x=1 (normal -x)
    lines += ["usemtl light", "f 9 10 11 12"]         # light quad facing down
    mtl = ["newmtl white", "Kd 0.725 0.71 0.68", "newmtl red", "Kd 0.63 0.065 0.05",
           "newmtl green", "Kd 0.14 0.45 0.091", "newmtl light", "Kd 1 1 1", "Ke 17 12 4"]
    return "\n".join(lines) + "\n", "\n".join(mtl) + "\n"


def block():
    lines = ["# unit cube, relative indices, no normals, no materials"]
    for z in (-0.5, 0.5):
        for y in (-0.5, 0.5):
            for x in (-0.5, 0.5):
                lines.append("v %r %r %r" % (x, y, z))
    # after 8 vertices, index -8 is vertex 1 (x-,y-,z-), ... -1 is vertex 8
    idx = lambda i: str(i - 9)  # noqa: E731  (absolute 1..8 -> relative)
    faces = [(1, 3, 4, 2), (5, 6, 8, 7), (1, 2, 6, 5), (3, 7, 8, 4), (1, 5, 7, 3), (2, 4, 8, 6)]
    for f in faces:
        lines.append("f " + " ".join(idx(i) for i in f))
    return "\n".join(lines) + "\n"


def octa():
    v = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    lines = ["# octahedron, explicit normals on the upper half"]
    lines += ["v %r %r %r" % p for p in v]
    lines += ["vn %r %r %r" % p for p in v]
    lines += ["usemtl gold"]
    up = [(1, 3, 5), (5, 3, 2), (2, 3, 6), (6, 3, 1)]
    down = [(5, 4, 1), (2, 4, 5), (6, 4, 2), (1, 4, 6)]
    for a, b, c in up:
        lines.append("f %d//%d %d//%d %d//%d" % (a, a, b, b, c, c))
    for a, b, c in down:
        lines.append("f %d/%d %d %d" % (a, a, b, c))  # v/vt form: texture index ignored, no normal
    return "\n".join(lines) + "\n"


def scene():
    return {
        "models": [
            {"name": "room", "path": "room.obj"},
            {"name": "block", "path": "block.obj"},
            {"name": "octa", "path": "octa.obj"},
            {"name": "unused", "path": "does_not_exist.obj"},
        ],
        "instances": [
            {"name": "room", "model": "room",
             "transform": [{"matrix": [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]}]},
            {"name": "tall", "model": "block",
             "transform": [{"scale": [0.3, 0.6, 0.3]}, {"rotate-y": 17.5}, {"translate": [0.33, 0.3, 0.65]}]},
            {"name": "short", "model": "block",
             "transform": [{"scale": 0.3}, {"rotate-y": -18.0}, {"translate": [0.66, 0.15, 0.35]}]},
            {"name": "gem", "model": "octa",
             "transform": [{"scale": 0.09}, {"rotate-x": 30.0}, {"rotate-z": 10.0}, {"translate": [0.5, 0.62, 0.45]}]},
            {"name": "ghost", "model": "missing"},
        ],
        "materials": [
            {"model": "block", "materials": [{"name": "DefaultMaterial", "diffuse": [0.7, 0.7, 0.7],
                                              "roughness": 0.6, "specular": 0.5}]},
            {"model": "octa", "materials": [{"name": "gold", "diffuse": [1.0, 0.78, 0.34], "metallic": 1.0,
                                             "roughness": 0.25, "clearcoat": 0.3, "diffuseMap": "gold.png"}]},
            {"model": "room", "materials": [{"name": "white", "roughness": 0.9},
                                            {"name": "light", "emissive": [17.0, 12.0, 4.0], "diffuse": [1, 1, 1]}]},
        ],
        "camera": {"position": [0.5, 0.5, -1.4], "target": [0.5, 0.5, 0.5], "up": [0.0, 1.0, 0.0], "fov": 40.0},
        "renderer": {"multisample": 4, "maxdepth": 5, "ray-epsilon": 0.001},
        "sky": {"enabled": False, "drawsun": True, "dir": [0.0, 1.0, 0.0], "angulardiameter": 32.0,
                "intensity": 1000.0, "turbidity": 4.0},
    }


def main():
    os.makedirs(HERE, exist_ok=True)
    obj, mtl = room()
    files = {"room.obj": obj, "room.mtl": mtl, "block.obj": block(), "octa.obj": octa(),
             "box_scene.json": json.dumps(scene(), indent=2) + "\n"}
    for name, text in files.items():
        with open(os.path.join(HERE, name), "w") as f:
            f.write(text)
    print("wrote", ", ".join(sorted(files)))


if __name__ == "__main__":
    main()
