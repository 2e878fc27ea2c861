"""Scene descriptions shared by the oracle and the library in the tests.

cornell34: the Cornell box of the reference's example scene
(example/cornell-box-specular.pbrt:1-91) with its sphere and rotated block
replaced by two axis-aligned boxes: 5 wall quads (10 triangles), a grey diffuse
short box and a specular silver (metal-Ag, roughness 0) tall box (12 triangles
each), and the example's emissive quad (rgb L 1, scale 20, 2 triangles).
"""
import numpy as np

GREY, RED, GREEN = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)


def quad(a, b, c, d):
    """Triangles of indices [0 1 2 2 3 0] over the quad a b c d."""
    return [a + b + c, c + d + a]


def box(x0, x1, y0, y1, z0, z1):
    p = lambda x, y, z: (x, y, z)
    faces = [
        (p(x0, y0, z0), p(x1, y0, z0), p(x1, y0, z1), p(x0, y0, z1)),  # bottom
        (p(x0, y1, z0), p(x0, y1, z1), p(x1, y1, z1), p(x1, y1, z0)),  # top
        (p(x0, y0, z0), p(x0, y1, z0), p(x1, y1, z0), p(x1, y0, z0)),  # front (z0)
        (p(x0, y0, z1), p(x1, y0, z1), p(x1, y1, z1), p(x0, y1, z1)),  # back (z1)
        (p(x0, y0, z0), p(x0, y0, z1), p(x0, y1, z1), p(x0, y1, z0)),  # left (x0)
        (p(x1, y0, z0), p(x1, y1, z0), p(x1, y1, z1), p(x1, y0, z1)),  # right (x1)
    ]
    out = []
    for f in faces:
        out += quad(*f)
    return out


def cornell34(width=256, height=256, spp=4, max_depth=5):
    tris, mats, lights = [], [], []
    # materials: 0 grey, 1 red, 2 green, 3 silver
    materials = [
        {"type": "diffuse", "rgb": GREY},
        {"type": "diffuse", "rgb": RED},
        {"type": "diffuse", "rgb": GREEN},
        {"type": "conductor", "metal": "Ag", "roughness": 0.0, "remap": True},
    ]

    def add(ts, m, light=-1):
        for t in ts:
            tris.append(t)
            mats.append(m)
            lights.append(light)

    # light (example :18-29), grey material, emitter 0
    add(quad((343.0, 548.7, 227.0), (343.0, 548.7, 332.0), (213.0, 548.7, 332.0), (213.0, 548.7, 227.0)), 0, 0)
    add(quad((552.8, 0.0, 0.0), (549.6, 0.0, 559.2), (556.0, 548.8, 559.2), (556.0, 548.8, 0.0)), 1)  # red
    add(quad((0.0, 0.0, 559.2), (0.0, 0.0, 0.0), (0.0, 548.8, 0.0), (0.0, 548.8, 559.2)), 2)  # green
    add(quad((552.8, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 559.2), (549.6, 0.0, 559.2)), 0)  # floor
    add(quad((556.0, 548.8, 0.0), (556.0, 548.8, 559.2), (0.0, 548.8, 559.2), (0.0, 548.8, 0.0)), 0)  # ceiling
    add(quad((549.6, 0.0, 559.2), (0.0, 0.0, 559.2), (0.0, 548.8, 559.2), (556.0, 548.8, 559.2)), 0)  # back
    add(box(130.0, 290.0, 0.0, 165.0, 65.0, 225.0), 0)  # short box, grey
    add(box(265.0, 430.0, 0.0, 330.0, 270.0, 435.0), 3)  # tall box, silver
    return {
        "eye": (278.0, 273.0, -800.0), "look": (278.0, 273.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 37.5,
        "width": width, "height": height, "spp": spp, "max_depth": max_depth,
        "tri": np.array(tris, np.float32).reshape(-1, 9),
        "tri_material": np.array(mats, np.uint32), "tri_light": np.array(lights, np.int32),
        "materials": materials, "lights": [{"rgb": (1.0, 1.0, 1.0), "scale": 20.0}],
    }
