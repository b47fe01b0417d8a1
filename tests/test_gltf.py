"""glTF 2.0 models as scenes (rt_scene_build with a FILE.gltf or FILE.glb name), on the CPU: small files written here
(JSON with base64 data: buffers, a .glb built in memory, a .gltf whose .bin lies beside it) load to exactly the
triangles, indices and node transforms they hold; the loader refuses what the reference's loader refuses, with its
reasons, including the reference's own fixtures (tests/data/cut-short.glb: "the file is shorter than it says";
tests/data/lines-only.gltf: a model without a triangle, warned that points and lines are skipped, which the command
line refuses); the metallic-roughness factors map to the renderer's materials."""
import base64
import ctypes as C
import json
import math
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "gltf")


class Writer:
    """A glTF 2.0 file assembled from parts: one buffer (a data: URI, a .glb BIN chunk or a file beside the model)"""

    def __init__(self):
        self.buffer = bytearray()
        self.views, self.accessors, self.meshes, self.nodes, self.materials = [], [], [], [], []
        self.extra = {}

    def _view(self, data: bytes) -> int:
        while len(self.buffer) % 4:
            self.buffer.append(0)
        self.views.append({"buffer": 0, "byteOffset": len(self.buffer), "byteLength": len(data)})
        self.buffer += data
        return len(self.views) - 1

    def positions(self, pts) -> int:
        a = np.asarray(pts, np.float32).reshape(-1, 3)
        v = self._view(a.tobytes())
        self.accessors.append({"bufferView": v, "componentType": 5126, "count": len(a), "type": "VEC3",
                               "min": a.min(0).tolist(), "max": a.max(0).tolist()})
        return len(self.accessors) - 1

    def indices(self, idx, dtype=np.uint32) -> int:
        a = np.asarray(idx, dtype)
        v = self._view(a.tobytes())
        ctype = {np.uint8: 5121, np.uint16: 5123, np.uint32: 5125}[dtype]
        self.accessors.append({"bufferView": v, "componentType": ctype, "count": len(a), "type": "SCALAR"})
        return len(self.accessors) - 1

    def mesh(self, *primitives) -> int:
        self.meshes.append({"primitives": list(primitives)})
        return len(self.meshes) - 1

    def node(self, **n) -> int:
        self.nodes.append(n)
        return len(self.nodes) - 1

    def text(self, uri=None, scene_nodes=None) -> dict:
        buf = {"byteLength": len(self.buffer)}
        if uri is not None:
            buf["uri"] = uri
        doc = {"asset": {"version": "2.0"}, "scene": 0,
               "scenes": [{"nodes": scene_nodes if scene_nodes is not None else [0]}],
               "nodes": self.nodes, "meshes": self.meshes, "accessors": self.accessors, "bufferViews": self.views,
               "buffers": [buf]}
        if self.materials:
            doc["materials"] = self.materials
        doc.update(self.extra)
        return doc

    def save(self, path, scene_nodes=None):
        if path.endswith(".glb"):
            js = json.dumps(self.text(None, scene_nodes)).encode()
            js += b" " * (-len(js) % 4)
            bin_ = bytes(self.buffer) + b"\0" * (-len(self.buffer) % 4)
            body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bin_), 0x004E4942) + bin_
            data = struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body
            with open(path, "wb") as f:
                f.write(data)
        else:
            uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.buffer)).decode()
            with open(path, "w") as f:
                json.dump(self.text(uri, scene_nodes), f)
        return path


TRIANGLE = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
SQUARE = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]


def meshes_of(scene):
    d = scene.desc
    out = []
    for k in range(d.mesh_count):
        m = d.meshes[k]
        pos = np.ctypeslib.as_array(m.positions, (m.vertex_count * 3,)).reshape(-1, 3).copy()
        idx = np.ctypeslib.as_array(m.indices, (m.triangle_count * 3,)).reshape(-1, 3).copy()
        out.append((pos, idx, m.material))
    inst = [(d.instances[k].mesh, np.array(d.instances[k].object_to_world[:]).reshape(3, 4))
            for k in range(d.instance_count)]
    return out, inst


def test_one_triangle(rt, tmp_path):
    w = Writer()
    w.mesh({"attributes": {"POSITION": w.positions(TRIANGLE)}})
    w.node(mesh=0)
    s = rt.build_scene(w.save(str(tmp_path / "flat.gltf")), width=64)
    meshes, inst = meshes_of(s)
    assert len(meshes) == 1 and np.array_equal(meshes[0][0], np.float32(TRIANGLE))
    assert np.array_equal(meshes[0][1], [[0, 1, 2]])
    assert len(inst) == 1 and np.array_equal(inst[0][1], np.eye(3, 4))
    assert s.warnings == []
    # glTF's default material (metallic 1, roughness 1): a GGX conductor of base color 1
    d = s.desc
    mat = d.materials[meshes[0][2]]
    assert mat.kind == rt.MATERIAL_KINDS["metal"] and (mat.albedo.x, mat.albedo.y, mat.albedo.z) == (1, 1, 1)
    assert d.microfacets[meshes[0][2]].kind == 1 and d.microfacets[meshes[0][2]].roughness == 1


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, None])
def test_square_with_each_index_type(rt, tmp_path, dtype):
    w = Writer()
    prim = {"attributes": {"POSITION": w.positions(SQUARE if dtype else [SQUARE[i] for i in (0, 1, 2, 0, 2, 3)])}}
    if dtype:
        prim["indices"] = w.indices([0, 1, 2, 0, 2, 3], dtype)
    w.mesh(prim)
    w.node(mesh=0)
    meshes, _ = meshes_of(rt.build_scene(w.save(str(tmp_path / "square.gltf")), width=64))
    pos, idx, _ = meshes[0]
    assert len(idx) == 2
    tri = pos[idx]
    expect = np.float32([[SQUARE[0], SQUARE[1], SQUARE[2]], [SQUARE[0], SQUARE[2], SQUARE[3]]])
    assert np.array_equal(tri, expect)


def test_strips_and_fans(rt, tmp_path):
    w = Writer()
    p = w.positions(SQUARE)
    w.mesh({"attributes": {"POSITION": p}, "indices": w.indices([0, 1, 3, 2]), "mode": 5},
           {"attributes": {"POSITION": p}, "indices": w.indices([0, 1, 2, 3]), "mode": 6})
    w.node(mesh=0)
    meshes, inst = meshes_of(rt.build_scene(w.save(str(tmp_path / "strip.gltf")), width=64))
    assert [m[1].tolist() for m in meshes] == [[[0, 1, 3], [3, 1, 2]], [[0, 1, 2], [0, 2, 3]]]
    assert [i[0] for i in inst] == [0, 1]


def test_node_transforms(rt, tmp_path):
    """translation * rotation * scale, a matrix node, and a parent's transform applied to its child's"""
    w = Writer()
    w.mesh({"attributes": {"POSITION": w.positions(TRIANGLE)}})
    h = math.sqrt(0.5)
    w.node(translation=[1, 2, 3], children=[1, 2])
    w.node(mesh=0, rotation=[0, h, 0, h], scale=[2, 2, 2])  # 90 degrees about y
    w.node(mesh=0, matrix=[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 5, 0, 0, 1])
    _, inst = meshes_of(rt.build_scene(w.save(str(tmp_path / "nodes.gltf")), width=64))
    assert len(inst) == 2
    rot = np.array([[0, 0, 2, 1], [0, 2, 0, 2], [-2, 0, 0, 3]], float)
    assert np.allclose(inst[0][1], rot, atol=1e-12)
    assert np.array_equal(inst[1][1], [[1, 0, 0, 6], [0, 1, 0, 2], [0, 0, 1, 3]])


def test_glb_and_buffer_file(rt, tmp_path):
    w = Writer()
    w.mesh({"attributes": {"POSITION": w.positions(SQUARE)}, "indices": w.indices([0, 1, 2, 0, 2, 3], np.uint16)})
    w.node(mesh=0)
    glb = meshes_of(rt.build_scene(w.save(str(tmp_path / "model.glb")), width=64))
    with open(tmp_path / "the data.bin", "wb") as f:
        f.write(bytes(w.buffer))
    with open(tmp_path / "beside.gltf", "w") as f:
        json.dump(w.text("the%20data.bin"), f)
    beside = meshes_of(rt.build_scene(str(tmp_path / "beside.gltf"), width=64))
    assert np.array_equal(glb[0][0][0], beside[0][0][0]) and np.array_equal(glb[0][0][1], beside[0][0][1])
    assert np.array_equal(glb[0][0][1], [[0, 1, 2], [0, 2, 3]])


def test_a_mesh_on_two_nodes_is_shared(rt, tmp_path):
    w = Writer()
    w.mesh({"attributes": {"POSITION": w.positions(TRIANGLE)}})
    w.node(mesh=0, translation=[-1, 0, 0])
    w.node(mesh=0, translation=[1, 0, 0])
    meshes, inst = meshes_of(rt.build_scene(w.save(str(tmp_path / "two.gltf"), scene_nodes=[0, 1]), width=64))
    assert len(meshes) == 1 and [i[0] for i in inst] == [0, 0]
    assert [i[1][0, 3] for i in inst] == [-1, 1]


def test_materials(rt, tmp_path):
    w = Writer()
    p = w.positions(TRIANGLE)
    w.materials = [{"pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.6, 0.3, 1], "metallicFactor": 1,
                                             "roughnessFactor": 0.25}},
                   {"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.8, 1], "metallicFactor": 0,
                                             "roughnessFactor": 0.5, "baseColorTexture": {"index": 0}}},
                   {"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.5, 0.5, 1], "metallicFactor": 0}}]
    w.mesh(*[{"attributes": {"POSITION": p}, "material": k} for k in range(3)])
    w.node(mesh=0)
    s = rt.build_scene(w.save(str(tmp_path / "mats.gltf")), width=64)
    meshes, _ = meshes_of(s)
    d = s.desc
    kinds = [(d.materials[m[2]].kind, d.microfacets[m[2]].kind, d.microfacets[m[2]].roughness) for m in meshes]
    assert kinds == [(rt.MATERIAL_KINDS["metal"], 1, 0.25), (rt.MATERIAL_KINDS["lambertian"], 2, 0.5),
                     (rt.MATERIAL_KINDS["lambertian"], 0, 0.0)]
    assert s.warnings == ["textures are not rendered: the materials' factors are used"]


def test_studio_frames_the_model(rt, tmp_path):
    w = Writer()
    w.mesh({"attributes": {"POSITION": w.positions(np.array(TRIANGLE) * 10 + [100, 50, 0])}})
    w.node(mesh=0)
    s = rt.build_scene(w.save(str(tmp_path / "far.gltf")), width=160, samples=4)
    cam = s.desc.camera
    assert (cam.image_width, cam.image_height) == (160, 90)
    # the camera looks at the model's bounds: their center projects near the image's center
    center = np.array([105, 55, 0], float)
    eye = np.array([cam.center.x, cam.center.y, cam.center.z])
    p00 = np.array([cam.pixel00_loc.x, cam.pixel00_loc.y, cam.pixel00_loc.z])
    du = np.array([cam.pixel_delta_u.x, cam.pixel_delta_u.y, cam.pixel_delta_u.z])
    dv = np.array([cam.pixel_delta_v.x, cam.pixel_delta_v.y, cam.pixel_delta_v.z])
    mid = p00 + 79.5 * du + 44.5 * dv
    a, b = mid - eye, center - eye
    assert np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b) > 0.999
    assert s.desc.environment and s.desc.quad_count == 1  # the sky and the ground


def test_reference_fixtures(rt):
    with pytest.raises(RuntimeError, match="the file is shorter than it says"):
        rt.build_scene(os.path.join(GOLDEN, "cut-short.glb"))
    with pytest.raises(ValueError, match="points and lines skipped"):
        rt.build_scene(os.path.join(GOLDEN, "lines-only.gltf"))


def test_refusals(rt, tmp_path):
    def model(**changes):
        w = Writer()
        w.mesh({"attributes": {"POSITION": w.positions(TRIANGLE)}, "indices": w.indices(changes.pop("idx", [0, 1, 2]))})
        w.node(mesh=0, **changes.pop("node", {}))
        for k, v in changes.items():
            w.extra[k] = v
        return w

    with pytest.raises(RuntimeError, match="an index past the vertices"):
        rt.build_scene(model(idx=[0, 1, 3]).save(str(tmp_path / "index.gltf")))
    with pytest.raises(RuntimeError, match="cycle"):
        w = model(node={"children": [0]})
        rt.build_scene(w.save(str(tmp_path / "cycle.gltf")))
    with pytest.raises(RuntimeError, match="it requires KHR_draco_mesh_compression"):
        rt.build_scene(model(extensionsRequired=["KHR_draco_mesh_compression"],
                             extensionsUsed=["KHR_draco_mesh_compression"]).save(str(tmp_path / "draco.gltf")))
    w = model()
    w.accessors[0]["count"] = 4  # four positions in a view of three
    with pytest.raises(RuntimeError, match="past its buffer view"):
        rt.build_scene(w.save(str(tmp_path / "short.gltf")))
    w = model()
    with open(tmp_path / "v1.gltf", "w") as f:
        doc = w.text("data:application/octet-stream;base64," + base64.b64encode(bytes(w.buffer)).decode())
        doc["asset"]["version"] = "1.0"
        json.dump(doc, f)
    with pytest.raises(RuntimeError, match="2.0"):
        rt.build_scene(str(tmp_path / "v1.gltf"))
    with open(tmp_path / "broken.gltf", "w") as f:
        f.write('{"asset": {"version": "2.0"}, ')
    with pytest.raises(RuntimeError, match="doesn't parse"):
        rt.build_scene(str(tmp_path / "broken.gltf"))
    with pytest.raises(RuntimeError, match="can't open"):
        rt.build_scene(str(tmp_path / "nowhere.glb"))


def test_crafted_bounds_are_refused(rt, tmp_path):
    """Accessor and view bounds whose sums or products would wrap modulo 2^64 are refused cleanly, before any read
    (count 32769 with byteStride 2^49 passed a summed check and read far past the buffer)"""
    def model():
        w = Writer()
        w.mesh({"attributes": {"POSITION": w.positions(TRIANGLE)}, "indices": w.indices([0, 1, 2])})
        w.node(mesh=0)
        return w

    w = model()
    w.accessors[0]["count"] = 32769
    w.views[0]["byteStride"] = 2 ** 49
    with pytest.raises(RuntimeError, match="byteStride"):
        rt.build_scene(w.save(str(tmp_path / "wrap.gltf")))
    w = model()
    w.views[0]["byteStride"] = 2  # below glTF's minimum of 4
    with pytest.raises(RuntimeError, match="byteStride"):
        rt.build_scene(w.save(str(tmp_path / "stride2.gltf")))
    w = model()
    w.views[0]["byteStride"] = 8  # within [4, 252] but smaller than a VEC3 of floats
    with pytest.raises(RuntimeError, match="smaller than"):
        rt.build_scene(w.save(str(tmp_path / "stride8.gltf")))
    w = model()
    w.accessors[0]["byteOffset"] = 10 ** 15  # far past the view
    with pytest.raises(RuntimeError, match="past its buffer view"):
        rt.build_scene(w.save(str(tmp_path / "offset.gltf")))
    w = model()
    w.views[0]["byteOffset"] = 10 ** 15  # far past the buffer
    with pytest.raises(RuntimeError, match="past its buffer"):
        rt.build_scene(w.save(str(tmp_path / "view.gltf")))
    w = model()
    w.accessors[0]["count"] = 10 ** 12
    with pytest.raises(RuntimeError, match="elements"):
        rt.build_scene(w.save(str(tmp_path / "count.gltf")))


@pytest.mark.gpu
def test_model_renders_alike_in_both_integrators(rt, tmp_path):
    """A model (an octahedron of two materials, placed twice) renders on the GPU: finite, lit, the same bytes from the
    megakernel and the wavefront integrator and from one seed twice"""
    w = Writer()
    pts = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    faces = [[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]]
    p = w.positions(pts)
    w.materials = [{"pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.5, 0.2, 1], "metallicFactor": 1,
                                             "roughnessFactor": 0.3}},
                   {"pbrMetallicRoughness": {"baseColorFactor": [0.3, 0.6, 0.9, 1], "metallicFactor": 0,
                                             "roughnessFactor": 0.6}}]
    w.mesh({"attributes": {"POSITION": p}, "indices": w.indices(np.ravel(faces[:4])), "material": 0},
           {"attributes": {"POSITION": p}, "indices": w.indices(np.ravel(faces[4:])), "material": 1})
    w.node(mesh=0, translation=[-1.2, 1, 0])
    w.node(mesh=0, translation=[1.2, 1, 0], rotation=[0, 0.3826834, 0, 0.9238795])
    s = rt.build_scene(w.save(str(tmp_path / "octa.glb"), scene_nodes=[0, 1]), width=96, samples=16)
    images = []
    for integrator in ("megakernel", "wavefront", "megakernel"):
        with rt.Session(s, seed=3, integrator=integrator) as sess:
            sess.render_to(s.passes)
            images.append(sess.pixels())
    assert np.isfinite(images[0]).all() and images[0].mean() > 0.05
    assert np.array_equal(images[0], images[1]) and np.array_equal(images[0], images[2])
