"""Scene, material and frame files (include/ff_io.h) against what the
reference's own import_export code returns on the same files
(tests/golden/io.npz, written by tests/golden/make_io_golden.py from the
reference sources compiled into oracle/_ref): scenes, materials, task
frames, exported text and CSV bytes must be identical. Host code only."""
import os

import numpy as np
import pytest

from conftest import bits, golden

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")

OBJ_CASES = ["parallel_planes.obj", "perpendicular_planes.obj", "test_material.obj", "unknown_material.obj",
             "tgrid_10x10.obj", "grid_10x10.obj", "confetti_2k.obj", "patch_sphere_2k.obj",
             "edge/formats.obj", "edge/crlf.obj", "edge/crlf_lib.obj", "edge/nogroup.obj", "edge/bad_vertex.obj",
             "edge/bad_face.obj", "edge/missing_mtl.obj", "edge/bad_value.obj", "edge/bad_newmtl.obj",
             "edge/trailing_space.obj", "edge/no_lib.obj", "edge/bad_usemtl.obj", "edge/usemtl_b.obj"]
TASK_CASES = [("parallel_planes.task", 2, 1), ("grid_10x10_heat.task", 100, 1), ("tgrid_10x10_heat.task", 200, 1),
              ("confetti_2k.task", 2000, 1), ("patch_sphere_2k_sun.task", 2000, 1), ("edge/frames.task", 3, 3),
              ("edge/frames.task", 2, 4), ("edge/bad_step.task", 1, 1), ("edge/bad_tmprt.task", 1, 1)]
REEXPORT = ["test_material.obj", "tgrid_10x10.obj", "edge/formats.obj", "edge/nogroup.obj"]


def key(name):
    return name.replace("/", "__").replace(".", "_")


@pytest.fixture(scope="module")
def io():
    from form_factors_amd import formats

    formats.lib()
    return formats


@pytest.fixture(scope="module")
def g():
    return golden("io.npz")


def names_of(arr):
    return [bytes(n).split(b"\0", 1)[0].decode("latin-1") for n in arr]


@pytest.mark.parametrize("name", OBJ_CASES)
def test_obj_import_matches_reference(io, g, name):
    k = key(name)
    want = int(g[f"obj_{k}_code"])
    try:
        got = io.obj_import(os.path.join(MODELS, name))
        code = 0
    except io.FormatError as e:
        got, code = None, e.code
    assert code == want, (name, code, want)
    if want == 0:
        assert np.array_equal(bits(got["faces"]), bits(g[f"obj_{k}_faces"]))
        assert np.array_equal(got["meshes"], g[f"obj_{k}_meshes"])
        assert np.array_equal(bits(got["materials"]), bits(g[f"obj_{k}_props"]))
        assert got["names"] == names_of(g[f"obj_{k}_names"])


def test_obj_import_rules(io, g):
    """What the golden results above exercise, spelled out: faces before the
    first group belong to no mesh, a quad keeps its first three vertices, an
    empty group has no faces, the material line after a block's fourteenth
    property is consumed (mat_b is lost), a duplicate property keeps the
    later value, CRLF names keep their '\\r'."""
    s = io.obj_import(os.path.join(MODELS, "edge", "formats.obj"))
    assert s["meshes"].tolist() == [[1, 3, 1], [4, 1, 2], [5, 0, 0], [5, 2, 0]]
    assert s["names"] == ["mat_a", "mat_c", "mat_f"]
    assert s["materials"][2, 0] == 9.0  # shell.density given twice
    quad = s["faces"][3].reshape(3, 3)
    assert np.array_equal(quad, np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1]], np.float32))  # f 5 6 8 7
    c = io.obj_import(os.path.join(MODELS, "edge", "crlf_lib.obj"))
    assert c["names"] == ["mat_d\r", "mat_e\r"] and c["meshes"].tolist() == [[0, 1, 1]]
    n = io.obj_import(os.path.join(MODELS, "edge", "nogroup.obj"))
    assert n["meshes"].tolist() == [[0, 2, 0]] and n["names"] == ["black_body"]
    assert n["materials"][0].tolist() == [1, 1, 1, 1, 0, 0, 1, 0, 1, 0, 0, 1, 0, 1]


@pytest.mark.parametrize("name", ["edge/face_range.obj", "edge/usemtl_first.obj"])
def test_obj_import_undefined_inputs_are_format_errors(io, name):
    """Inputs on which the reference's importer has undefined behaviour (a
    face index past the vertex list; usemtl before any group) are rejected."""
    with pytest.raises(io.FormatError) as e:
        io.obj_import(os.path.join(MODELS, name))
    assert e.value.code == -io.FORMAT_ERROR


def test_reference_import_tests(io):
    """The reference's own import tests (test/import.cpp:23-66)."""
    s = io.obj_import(os.path.join(MODELS, "parallel_planes.obj"))
    assert len(s["faces"]) == 4 and len(s["meshes"]) == 2
    with pytest.raises(io.FormatError) as e:
        io.obj_import(os.path.join(MODELS, "unknown_material.obj"))
    assert e.value.code == -io.MATERIAL_NOT_DEFINED
    m = io.obj_import(os.path.join(MODELS, "test_material.obj"))
    assert len(m["names"]) == 1
    assert np.allclose(m["materials"][0], [1, 2, 3, 4, 1.1, 1.2, 1.3, 1.4, 1.5, 2.1, 2.2, 2.3, 2.4, 2.5], atol=0.01)


@pytest.mark.parametrize("name,n,calls", TASK_CASES)
def test_task_import_matches_reference(io, g, name, n, calls):
    k = f"task_{key(name)}_{n}_{calls}"
    codes = g[k + "_codes"]
    if codes[0] < 0:
        with pytest.raises(io.FormatError) as e:
            io.task_import(os.path.join(MODELS, name), n, calls)
        assert e.value.code == codes[0]
        return
    got = io.task_import(os.path.join(MODELS, name), n, calls)
    for c, (count, dt, temps) in enumerate(got["frames"]):
        assert count == codes[c]
        assert np.float32(dt) == g[k + "_dts"][c]
        assert np.array_equal(bits(temps), bits(g[k + "_temps"][c, :count]))
    src = np.array(got["heat_sources"], np.float32).reshape(-1, 2)
    assert np.array_equal(bits(src), bits(g[k + "_src"]))
    dst = g[k + "_dst"]
    if len(dst):
        assert np.array_equal(bits(np.array(got["distant_source"], np.float32)), bits(dst))
    else:
        assert got["distant_source"] is None


@pytest.mark.parametrize("name", REEXPORT)
def test_scene_export_matches_reference(io, g, tmp_path, name):
    s = io.obj_import(os.path.join(MODELS, name))
    out = tmp_path / "scene.obj"
    io.obj_export_scene(str(out), s["faces"], s["meshes"], s["materials"], s["names"])
    assert out.read_bytes() == bytes(g[f"reexport_{key(name)}"])


@pytest.mark.parametrize("binary", [0, 1])
def test_frame_export_matches_reference(io, g, tmp_path, binary):
    out = tmp_path / "frames.txt"
    io.task_export(str(out), [(7, np.float32(0.01), g["frame_temps"])], binary=bool(binary))
    assert out.read_bytes() == bytes(g[f"frame_export_{binary}"])


def test_csv_export_matches_reference(io, g, tmp_path):
    out = tmp_path / "f.csv"
    io.csv_export(str(out), g["csv_F"])
    assert out.read_bytes() == bytes(g["csv_text"])
    # the reference's own CSV test (test/export.cpp:36-52)
    io.csv_export(str(out), np.array([[0, 1], [0, 1]], np.float32))
    rows = [list(map(float, line.split(";"))) for line in out.read_text().splitlines()]
    assert rows == [[0.0, 1.0], [0.0, 1.0]]
    with pytest.raises(io.FormatError) as e:
        io.csv_export(str(tmp_path / "no_such_dir" / "f.csv"), np.zeros((1, 1), np.float32))
    assert e.value.code == -io.CSV_FILE_ERROR


@pytest.mark.parametrize("name", ["test_material.obj", "tgrid_10x10.obj", "edge/formats.obj"])
def test_inline_materials_read_the_reference_export(io, g, tmp_path, name):
    """The reference's exporter writes material blocks into the OBJ file; the
    inline mode reads that text back to the scene the reference imported
    from the original files (vertices within the exporter's 6 decimals)."""
    k = key(name)
    path = tmp_path / "exported.obj"
    path.write_bytes(bytes(g[f"reexport_{k}"]))
    with pytest.raises(io.FormatError):  # the reference's importer cannot read it
        io.obj_import(str(path))
    s = io.obj_import(str(path), inline_materials=True)
    want_meshes = g[f"obj_{k}_meshes"]
    assert np.array_equal(s["names"], names_of(g[f"obj_{k}_names"]))
    assert np.array_equal(bits(s["materials"]), bits(g[f"obj_{k}_props"]))
    # only faces of meshes are written as f lines (a face outside every mesh
    # keeps its vertices but is not read back)
    assert np.array_equal(s["meshes"][:, 1:], want_meshes[:, 1:])
    faces = g[f"obj_{k}_faces"]
    want = np.concatenate([faces[a:a + n] for a, n, _ in want_meshes])
    assert s["faces"].shape == want.shape
    assert np.abs(s["faces"] - want).max() <= 5e-7 * max(1.0, np.abs(want).max())


def unpacked(name, tmp_path):
    """A model file under tests/golden/models; a .gz one is written out
    uncompressed into tmp_path first (the importers read files)."""
    path = os.path.join(MODELS, name)
    if not name.endswith(".gz"):
        return path
    import gzip

    out = tmp_path / os.path.basename(name[:-3])
    out.write_bytes(gzip.decompress(open(path, "rb").read()))
    return str(out)


@pytest.mark.parametrize("name,n_meshes", [("confetti_2k.obj", 2000), ("grid_10x10.obj", 100),
                                           ("patch_sphere_2k.obj", 2000), ("patch_sphere_20k.obj.gz", 20000)])
def test_inline_materials_read_the_reference_models(io, tmp_path, name, n_meshes):
    """The reference's models carry their materials inline: the inline mode
    reads them; vertices and faces equal a plain parse of the v / f lines.
    patch_sphere_20k.obj (models/, 20,000 one-face meshes) is kept gzipped."""
    path = unpacked(name, tmp_path)
    s = io.obj_import(path, inline_materials=True)
    verts, tris = [], []
    for line in open(path):
        if line.startswith("v "):
            verts.append([np.float32(x) for x in line.split()[1:4]])
        elif line.startswith("f "):
            tris.append([int(x.split("/")[0]) - 1 for x in line.split()[1:4]])
    faces = np.array(verts, np.float32)[np.array(tris)].reshape(-1, 9)
    assert np.array_equal(bits(s["faces"]), bits(faces))
    assert len(s["meshes"]) == n_meshes and s["meshes"][:, 1].sum() == len(faces)
    assert len(s["names"]) >= 1 and (s["meshes"][:, 2] < len(s["names"])).all()


def test_patch_sphere_20k_sun_task(io, tmp_path):
    """models/patch_sphere_20k_sun.task: one frame of 20,000 temperatures
    (three meshes at 400 K, the rest at 10 K), step 0.01 s and a 30 W/m^2
    distant source along -x; the scene's one inline material is the black
    body (emissivity and absorbance 1 on both sides, heat capacity 20)."""
    s = io.obj_import(unpacked("patch_sphere_20k.obj.gz", tmp_path), inline_materials=True)
    t = io.task_import(unpacked("patch_sphere_20k_sun.task.gz", tmp_path), 20000, 1, step=0.1)
    count, dt, T = t["frames"][0]
    assert count == 20000 and np.float32(dt) == np.float32(0.01)
    assert np.array_equal(np.flatnonzero(T == 400.0), [0, 1, 2]) and (T[3:] == 10.0).all()
    assert t["distant_source"] == (30.0, -1.0, 0.0, 0.0) and len(t["heat_sources"]) == 0
    assert s["names"] == ["black_body"]
    assert s["materials"][0].tolist() == [1, 20, np.float32(0.1), 1, 0, 0, 1, 0, 1, 0, 0, 1, 0, 1]
