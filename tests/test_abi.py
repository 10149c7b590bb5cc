"""The C ABI (include/convector.h) on the CPU: the library loads and exports
every declared entry point, and the host-side OBJ reader matches an
independent restatement of wavefront.rs with the reference's error cases."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import INDOOR_MATERIALS, MODEL_MATERIALS, read_obj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "convector.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cv_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(cv):
    names = _declared_functions()
    assert len(names) >= 24
    lib = ctypes.CDLL(cv._native.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(cv._native.EXPORTS)


def test_version_and_status_strings(cv):
    L = cv.lib()
    assert b"gfx950" in L.cv_version()
    assert L.cv_status_string(0) == b"ok"
    assert L.cv_status_string(7).startswith(b"fewer than 8")


@pytest.mark.parametrize("name", ["suzanne", "indoor", "box_windows", "suzannes_in_box", "plane",
                                  "stanford_bunny"])
def test_obj_loader_matches_restatement(cv, models_dir, name):
    mats = MODEL_MATERIALS.get(name, {})
    path = str(models_dir / (name + ".obj"))
    m = cv.Mesh.load_with_materials(path, mats)
    v, uv, mt = m.triangles()
    rv, ruv, rmt = read_obj(path, mats)
    assert v.shape == rv.shape
    assert v.view(np.uint32).tolist() == rv.view(np.uint32).tolist()
    assert (uv.view(np.uint32) == ruv.view(np.uint32)).all()
    assert (mt == rmt).all()


def test_obj_fan_triangulation(cv, tmp_path):
    p = tmp_path / "quad.obj"
    p.write_text("v 0 0 0\nv 2 0 0\nv 3 1 0\nv 1 3 0\nv -1 1 0\nf 1 2 3 4 5\n")
    v, _, mt = cv.Mesh.load(str(p)).triangles()
    # (i0, i1, i2), (i0, i2, i3), (i0, i3, i4) (wavefront.rs:133-140)
    assert v.reshape(-1, 3, 3)[:, :, :2].tolist() == [
        [[0, 0], [2, 0], [3, 1]], [[0, 0], [3, 1], [1, 3]], [[0, 0], [1, 3], [-1, 1]]]
    assert (mt == 0x00FFFFFF).all()  # SMaterial::white() default


def test_obj_tex_coords_only_when_all_three(cv, tmp_path):
    p = tmp_path / "uv.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.5 0.25\nvt 1 0\nvt 0 1\n"
                 "f 1/1 2/2 3/3\nf 1/1 2 3/3\n")
    _, uv, _ = cv.Mesh.load(str(p)).triangles()
    assert uv[0].tolist() == [0.5, 0.25, 1, 0, 0, 1]
    assert uv[1].tolist() == [0] * 6


@pytest.mark.parametrize("text,status", [
    ("v 0 0 0\nv 1 0 0\nv 2 0 0\nf 1 2 3\n", "CV_ERR_DEGENERATE_TRIANGLE"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl nope\nf 1 2 3\n", "CV_ERR_UNKNOWN_MATERIAL"),
    ("v 0 0 0\nv 1 0 0\nf 1 2 3\n", "CV_ERR_PARSE"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//1 2//1 3//1\n", "CV_ERR_PARSE"),
    ("v 0 0 x\n", "CV_ERR_PARSE"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n", "CV_ERR_PARSE"),
])
def test_obj_loader_errors(cv, tmp_path, text, status):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(cv.CvError) as e:
        cv.Mesh.load(str(p))
    assert e.value.name == status


def test_missing_file(cv, tmp_path):
    with pytest.raises(cv.CvError) as e:
        cv.Mesh.load(str(tmp_path / "missing.obj"))
    assert e.value.name == "CV_ERR_IO"


def test_mesh_from_triangles_rejects_degenerate(cv):
    t = np.array([[0, 0, 0, 1, 0, 0, 2, 0, 0]], np.float32)
    with pytest.raises(cv.CvError) as e:
        cv.Mesh.from_triangles(t, np.zeros(1, np.uint32))
    assert e.value.name == "CV_ERR_DEGENERATE_TRIANGLE"


def test_generated_scene_roundtrips_through_obj(cv, tmp_path):
    """The generator's OBJ (%.9g) reloads bit for bit (config-0 scene)."""
    from convector_amd import scenes
    sd = scenes.config0(subdivisions=2)
    table = scenes.write_obj(str(tmp_path / "c0.obj"), sd)
    v, _, mt = cv.Mesh.load_with_materials(str(tmp_path / "c0.obj"), table).triangles()
    assert (v.view(np.uint32) == sd.tris.view(np.uint32)).all()
    assert (mt == sd.materials).all()


def test_config_scene_sizes():
    from convector_amd import scenes
    assert scenes.config0().ntris == 20506
    sd = scenes.config0()
    assert (sd.materials == scenes.MAT_SKY).sum() == 8
    assert scenes.MAT_GLOSSY_WHITE == 0x10FFFFFF


def _header_fields(struct):
    """Field names of `typedef struct <struct> { ... }` in declaration order."""
    text = open(os.path.join(ROOT, "include", "convector.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct " + struct + r" \{(.*?)\}", text, flags=re.S).group(1)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):
            m = re.search(r"\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*(\[[^\]]*\])?\s*$", part.strip())
            names.append(m.group(1))
    return names


@pytest.mark.parametrize("struct,mirror", [
    ("cv_tuning", "Tuning"), ("cv_render_opts", "RenderOpts"), ("cv_job", "Job"),
    ("cv_request", "Request"), ("cv_server_stats", "ServerStats"),
    ("cv_render_stats", "RenderStats"), ("cv_bvh_info", "BvhInfo"),
    ("cv_scene_opts", "SceneOpts"), ("cv_material_entry", "MaterialEntry")])
def test_struct_mirrors_match_header(cv, struct, mirror):
    """The ctypes mirrors of the ABI's structs have the header's fields in
    the header's order (a field added on one side only shifts the rest)."""
    fields = [f[0] for f in getattr(cv._native, mirror)._fields_]
    assert fields == _header_fields(struct), struct


def test_host_buffers_are_checked_before_the_library_sees_them(cv):
    """A wrong array (dtype, layout, size) raises in the binding; the C ABI
    takes raw pointers and would read or write past it."""
    j = cv.Job(None, 64, 16, hdr=np.zeros(64 * 16 * 3, np.float32))
    assert j._ptr(j.hdr) == j.hdr.ctypes.data
    assert j._ptr(12345) == 12345  # a device pointer passes as is
    assert j._ptr(None) is None
    with pytest.raises(ValueError, match="elements"):
        j._ptr(np.zeros(64 * 16 * 3 - 1, np.float32))
    with pytest.raises(TypeError, match="dtype"):
        j._ptr(np.zeros(64 * 16 * 3, np.float64))
    with pytest.raises(ValueError, match="contiguous"):
        j._ptr(np.zeros(64 * 16 * 6, np.float32)[::2])
    ro = np.zeros(64 * 16 * 3, np.float32)
    ro.flags.writeable = False
    with pytest.raises(ValueError, match="writeable"):
        j._ptr(ro)
    with pytest.raises(ValueError, match="elements"):
        j._ptr(np.zeros(64 * 16 * 7, np.float32), np.float32, 8, "features")
    with pytest.raises(TypeError, match="dtype"):
        j._ptr(np.zeros(64 * 16, np.int32), np.uint32, 1, "gbuffer")
