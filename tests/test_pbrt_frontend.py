"""The .pbrt front end (pbrtm_pbrt_parse), host only: the fixture parses to the
scene the tests build in Python, transforms compose as the reference's
graphics state does (CTM = CTM * T), and unsupported input is an error, not a
silent approximation."""
import ctypes
import math
import os

import numpy as np
import pytest

import pbrtm
import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "fixtures", "cornell34.pbrt")


def test_fixture_matches_python_scene():
    f = pbrtm.PbrtFile(FIXTURE)
    d, ref = f.to_dict(), scenes.cornell34()
    assert (f.spp, f.max_depth, f.output) == (4, 5, "cornell34.png")
    for k in ("eye", "look", "up"):
        assert d[k] == tuple(np.float32(ref[k]))
    assert d["fov"] == np.float32(ref["fov"]) and (d["width"], d["height"]) == (256, 256)
    assert np.array_equal(d["tri"].view(np.uint32), ref["tri"].view(np.uint32))
    assert np.array_equal(d["tri_material"], ref["tri_material"])
    assert np.array_equal(d["tri_light"] >= 0, ref["tri_light"] >= 0)
    assert len(d["lights"]) == 2 and d["lights"][0] == {"rgb": (1.0, 1.0, 1.0), "scale": 20.0, "two_sided": False}
    for a, b in zip(d["materials"], ref["materials"]):
        assert a["type"] == b["type"]
        if a["type"] == "diffuse":
            assert a["rgb"] == tuple(float(np.float32(x)) for x in b["rgb"])
        else:
            assert a["metal"] == b["metal"] and a["uroughness"] == 0 and a["remap"]
    assert "xforms" not in d


def write(tmp_path, body, head=None):
    head = head if head is not None else 'Film "rgb" "integer xresolution" [8] "integer yresolution" [6]\nWorldBegin\n'
    p = tmp_path / "s.pbrt"
    p.write_text(head + body)
    return str(p)


TRI = 'Shape "trianglemesh" "integer indices" [0 1 2] "point3 P" [0 0 0 1 0 0 0 1 0]\n'


def test_transform_stack(tmp_path):
    body = ("AttributeBegin\nTranslate 1 2 3\nScale 2 4 8\n" + TRI + "AttributeEnd\n"
            "AttributeBegin\nRotate 90 0 0 1\n" + TRI + "AttributeEnd\n" + TRI +
            "AttributeBegin\nTransform [1 0 0 0  0 1 0 0  0 0 1 0  5 6 7 1]\nReverseOrientation\n" + TRI +
            "AttributeEnd\n")
    d = pbrtm.PbrtFile(write(tmp_path, body)).to_dict()
    assert list(d["tri_xform"]) == [0, 1, 0xFFFFFFFF, 2]
    assert list(d["tri_reverse"]) == [0, 0, 0, 1]
    ts = np.array([[2, 0, 0, 1], [0, 4, 0, 2], [0, 0, 8, 3], [0, 0, 0, 1]], np.float32)
    assert np.array_equal(d["xforms"][0], ts)
    rz = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    assert np.abs(d["xforms"][1] - rz).max() < 1e-7
    tr = np.eye(4, dtype=np.float32)
    tr[:3, 3] = (5, 6, 7)
    assert np.array_equal(d["xforms"][2], tr)


def test_named_materials_lights_and_defaults(tmp_path):
    body = ('MakeNamedMaterial "gold" "string type" "conductor" "spectrum eta" "metal-Au-eta" '
            '"spectrum k" "metal-Au-k" "float uroughness" 0.1 "float vroughness" 0.2 "bool remaproughness" false\n'
            + TRI + 'NamedMaterial "gold"\n' + TRI +
            'AttributeBegin\nMaterial "dielectric" "spectrum eta" "glass-BK7"\n'
            'AreaLightSource "diffuse" "rgb L" [0.5 0.25 1] "bool twosided" true\n' + TRI + 'AttributeEnd\n')
    d = pbrtm.PbrtFile(write(tmp_path, body)).to_dict()
    mats = d["materials"]
    assert mats[d["tri_material"][0]] == {"type": "diffuse", "rgb": (0.5, 0.5, 0.5)}
    g = mats[d["tri_material"][1]]
    assert g["metal"] == "Au" and not g["remap"] and abs(g["uroughness"] - 0.1) < 1e-7
    assert mats[d["tri_material"][2]]["eta"] == "glass-BK7"
    assert list(d["tri_light"]) == [-1, -1, 0]
    assert d["lights"][0]["two_sided"] and d["lights"][0]["scale"] == 1.0


@pytest.mark.parametrize("body,msg", [
    ('Shape "cylinder" "float radius" 1\n', "cylinder"),
    ('ObjectBegin "a"\nShape "sphere" "float radius" 1\nObjectEnd\n', "ObjectBegin"),
    ('ObjectInstance "missing"\n', "not found"),
    ('Shape "sphere" "float radius" 1 "float thetamax" 2\n', "thetamax"),
    ('Material "coateddiffuse"\n', "coateddiffuse"),
    ('LightSource "point"\n', "LightSource"),
    ('LightSource "infinite" "string filename" "sky.exr"\n', "image infinite"),
    ('Shape "trianglemesh" "integer indices" [0 1 2] "point3 P" [0 0 0 1 0 0 0 1 0] "normal N" [0 0 1 0 0 1 0 0 1]\n',
     "normal N"),
    ('Shape "trianglemesh" "integer indices" [0 1 5] "point3 P" [0 0 0 1 0 0 0 1 0]\n', "out of range"),
])
def test_unsupported_is_an_error(tmp_path, body, msg):
    with pytest.raises(pbrtm.PbrtmError, match=msg):
        pbrtm.PbrtFile(write(tmp_path, body))


def test_camera_must_be_a_lookat(tmp_path):
    head = 'Scale -1 1 1\nLookAt 0 0 0 0 0 1 0 1 0\nCamera "perspective"\nFilm "rgb"\nWorldBegin\n'
    with pytest.raises(pbrtm.PbrtmError, match="LookAt"):
        pbrtm.PbrtFile(write(tmp_path, TRI, head))
    head = 'Sampler "stratified" "integer pixelsamples" 10\nFilm "rgb"\nWorldBegin\n'
    f = pbrtm.PbrtFile(write(tmp_path, TRI, head))
    assert f.spp == 9  # rounded down to a square, as the stratified sampler needs


@pytest.mark.parametrize("name,flags", [("path", 0x00), ("volpath", 0x00), ("megakernelpath", 0x10),
                                        ("ambientocclusion", 0x20), ("surfacenormal", 0x30)])
def test_integrator_names(tmp_path, name, flags):
    """scene_builder.cu:289-313: "volpath" is the wavefront "path"; the megakernel
    integrators by their MegakernelIntegrator::create names."""
    lib = pbrtm.load()
    out = ctypes.c_uint32(0xFFFF)
    assert lib.pbrtm_integrator_from_name(name.encode(), ctypes.byref(out)) == pbrtm.OK and out.value == flags
    head = f'Integrator "{name}" "integer maxdepth" [3]\nFilm "rgb"\nWorldBegin\n'
    f = pbrtm.PbrtFile(write(tmp_path, TRI, head))
    assert pbrtm.INTEGRATORS[f.integrator] == flags and f.max_depth == 3


@pytest.mark.parametrize("name", ["bdpt", "mlt", "Path", ""])
def test_unsupported_integrators(tmp_path, name):
    lib = pbrtm.load()
    out = ctypes.c_uint32(0)
    assert lib.pbrtm_integrator_from_name(name.encode(), ctypes.byref(out)) == pbrtm.ERR_UNSUPPORTED
    if name:
        with pytest.raises(pbrtm.PbrtmError, match="integrator"):
            pbrtm.PbrtFile(write(tmp_path, TRI, f'Integrator "{name}"\nFilm "rgb"\nWorldBegin\n'))


def test_infinite_light_place_in_light_order(tmp_path):
    """LightSource "infinite" joins the light list where it appears: after the
    emissive triangles created before it (scene_builder.cu:494, :637)."""
    body = ('AreaLightSource "diffuse" "rgb L" [1 1 1]\n' + TRI +
            'LightSource "infinite" "rgb L" [0.2 0.4 0.9] "float scale" 3\n' +
            'LightSource "infinite"\n')
    d = pbrtm.PbrtFile(write(tmp_path, body)).to_dict()
    inf = d["infinite_lights"]
    assert [l["light_index"] for l in inf] == [1, 1]
    assert inf[0]["rgb"] == tuple(np.float32([0.2, 0.4, 0.9])) and inf[0]["scale"] == 3 and inf[0]["illuminance"] < 0
    assert inf[1]["rgb"] is None and inf[1]["scale"] == 1
