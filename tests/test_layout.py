"""The plugin boundary's structs against the reference's own headers: a small
translation unit including the reference headers, one including
include/form_factors_amd/api.hpp and a C one including include/ff_plugin.h
print sizeof and the offset of every member; the three tables must agree, so
objects of the reference library pass through the plugin ABI unchanged.

Needs the reference tree (the build container); skipped elsewhere."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

REF_HEADERS = ["math/types.h", "ray_caster/system.h", "emission/system.h", "emission/malley_emission.h",
               "emission/parallel_rays_cpu.h", "form_factors/system.h", "subject/system.h",
               "thermal_equation/system.h", "thermal_equation/heat_source_cpu.h",
               "thermal_equation/parallel_rays_cpu.h", "thermal_equation/form_factors.h",
               "thermal_equation/radiance_cpu.h", "thermal_solution/system.h"]

# (reference type, api.hpp type, ff_plugin.h type or None, [(reference member, ff_plugin.h member)]);
# api.hpp members carry the reference's names. A None plugin member is not mirrored in C.
LAYOUTS = [
    ("math::vec3", "math::vec3", "ffp_vec3", [("x", "x"), ("y", "y"), ("z", "z")]),
    ("math::triangle_t", "math::triangle_t", "ffp_face", [("points", "points")]),
    ("math::ray_t", "math::ray_t", "ffp_ray", [("origin", "origin"), ("direction", "direction")]),
    ("ray_caster::task_t", "ray_caster::task_t", "ffp_rc_task",
     [("n_tasks", "n_tasks"), ("ray", "ray"), ("hit_face", "hit_face"), ("hit_point", "hit_point")]),
    ("ray_caster::scene_t", "ray_caster::scene_t", "ffp_rc_scene", [("n_faces", "n_faces"), ("faces", "faces")]),
    ("ray_caster::system_t", "ray_caster::system_t", None, [("methods", None)]),
    ("ray_caster::system_methods_t", "ray_caster::system_methods_t", None,
     [(m, None) for m in ("init", "shutdown", "set_scene", "prepare", "cast")]),
    ("emission::task_t", "emission::task_t", None, [("n_rays", None), ("rays", None)]),
    ("emission::system_methods_t", "emission::system_methods_t", None,
     [(m, None) for m in ("init", "shutdown", "set_scene", "calculate")]),
    ("malley_emission::task_t", "malley_emission::task_t", None,
     [("n_rays", None), ("rays", None), ("total_weight", None), ("weights", None)]),
    ("parallel_rays_emission_cpu::task_t", "parallel_rays_emission::task_t", None,
     [("n_rays", None), ("rays", None), ("direction", None), ("center", None), ("distance", None),
      ("radius", None)]),
    ("form_factors::mesh_t", "form_factors::mesh_t", "ffp_mesh",
     [("first_idx", "first_idx"), ("n_faces", "n_faces"), ("padding0", "padding0")]),
    ("form_factors::scene_t", "form_factors::scene_t", "ffp_ff_scene",
     [("n_faces", "n_faces"), ("faces", "faces"), ("n_meshes", "n_meshes"), ("meshes", "meshes")]),
    ("form_factors::task_t", "form_factors::task_t", "ffp_ff_task",
     [("n_rays", "n_rays"), ("form_factors", "form_factors")]),
    ("form_factors::system_methods_t", "form_factors::system_methods_t", None,
     [(m, None) for m in ("init", "shutdown", "set_scene", "prepare", "calculate")]),
    ("subject::shell_properties_t", "subject::shell_properties_t", None,
     [(m, None) for m in ("density", "heat_capacity", "thermal_conductivity", "thickness")]),
    ("subject::optical_properties_t", "subject::optical_properties_t", None,
     [(m, None) for m in ("specular_reflectance", "diffuse_reflectance", "absorbance", "transmittance",
                          "emissivity")]),
    ("subject::material_t", "subject::material_t", "ffp_material",
     [("shell.density", "density"), ("shell.heat_capacity", "heat_capacity"),
      ("shell.thermal_conductivity", "thermal_conductivity"), ("shell.thickness", "thickness"),
      ("front", "front"), ("front.emissivity", "front[4]"), ("rear", "rear"), ("rear.absorbance", "rear[2]"),
      ("name", "name")]),
    ("subject::mesh_t", "subject::mesh_t", "ffp_subject_mesh",
     [("first_idx", "first_idx"), ("n_faces", "n_faces"), ("material_idx", "material_idx")]),
    ("subject::scene_t", "subject::scene_t", "ffp_subject_scene",
     [(m, m) for m in ("n_faces", "faces", "n_meshes", "meshes", "n_materials", "materials")]),
    ("thermal_equation::task_t", "thermal_equation::task_t", "ffp_te_task",
     [("temperatures", "temperatures"), ("emission", "emission"), ("absorption", "absorption")]),
    ("thermal_equation::system_methods_t", "thermal_equation::system_methods_t", None,
     [(m, None) for m in ("init", "shutdown", "set_scene", "calculate")]),
    ("heat_source_equation::heat_source_t", "heat_source_equation::heat_source_t", "ffp_heat_source",
     [("mesh_idx", "mesh_idx"), ("power", "power")]),
    ("heat_source_equation::params_t", "heat_source_equation::params_t", None,
     [("n_sources", None), ("sources", None)]),
    ("parallel_rays_cpu::source_t", "parallel_rays_cpu::source_t", "ffp_source",
     [("power", "power"), ("direction", "direction")]),
    ("parallel_rays_cpu::params_t", "parallel_rays_cpu::params_t", None,
     [(m, None) for m in ("n_rays", "emitter", "source_param", "source")]),
    ("form_factors_equation::params_t", "form_factors_equation::params_t", None,
     [("form_factors_calculator", None), ("n_rays", None)]),
    ("radiance_equation::params_t", "radiance_equation::params_t", None, [("emitter", None), ("n_rays", None)]),
    ("thermal_solution::params_t", "thermal_solution::params_t", None,
     [("n_equations", None), ("equations", None)]),
    ("thermal_solution::task_t", "thermal_solution::task_t", "ffp_ts_task",
     [("n_step", "n_step"), ("time_delta", "time_delta"), ("temperatures", "temperatures")]),
    ("thermal_solution::system_methods_t", "thermal_solution::system_methods_t", None,
     [(m, None) for m in ("init", "shutdown", "set_scene", "calculate")]),
]

needs_ref = pytest.mark.skipif(not (os.path.exists(os.path.join(REF, "math", "types.h")) and shutil.which("g++")),
                               reason="reference headers or g++ not available here")


def _probe(tmp, name, includes, pick, lang, flags):
    lines = ["#include <stdio.h>", "#include <stddef.h>"] + [f'#include "{h}"' for h in includes]
    lines.append("int main(void) {")
    for k, entry in enumerate(LAYOUTS):
        typ, fields = pick(entry)
        if typ is None:
            continue
        lines.append(f'  printf("{k} size %zu\\n", sizeof({typ}));')
        for j, f in enumerate(fields):
            if f is not None:
                lines.append(f'  printf("{k} {j} %zu\\n", offsetof({typ}, {f}));')
    lines += ["  return 0;", "}"]
    src = os.path.join(tmp, f"{name}.{'c' if lang == 'c' else 'cpp'}")
    exe = os.path.join(tmp, name)
    with open(src, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    cc = "gcc" if lang == "c" else "g++"
    r = subprocess.run([cc, *flags, src, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.splitlines())


@needs_ref
def test_plugin_structs_match_the_reference_headers(tmp_path):
    tmp = str(tmp_path)
    ref = _probe(tmp, "ref", REF_HEADERS, lambda e: (e[0], [a for a, _ in e[3]]), "c++",
                 ["-std=c++11", "-fpermissive", "-w", "-Wno-invalid-offsetof", "-I", REF, "-include", "cstdio"])
    api = _probe(tmp, "api", ["form_factors_amd/api.hpp"], lambda e: (e[1], [a for a, _ in e[3]]), "c++",
                 ["-std=c++11", "-w", "-Wno-invalid-offsetof", "-I", os.path.join(ROOT, "include")])
    plug = _probe(tmp, "plugin", ["ff_plugin.h"], lambda e: (e[2], [b for _, b in e[3]]), "c",
                  ["-std=c99", "-w", "-I", os.path.join(ROOT, "include")])
    assert len(ref) == len(api) and len(ref) > 100
    for k, v in api.items():
        assert ref[k] == v, (LAYOUTS[int(k.split()[0])][0], k, ref[k], v)
    for k, v in plug.items():
        assert ref[k] == v, (LAYOUTS[int(k.split()[0])][2], k, ref[k], v)
    # the sizes the survey measured on x86-64
    size = {LAYOUTS[int(k.split()[0])][0]: int(v) for k, v in ref.items() if k.endswith("size")}
    assert size["math::triangle_t"] == 36 and size["math::ray_t"] == 24 and size["ray_caster::task_t"] == 32
    assert size["form_factors::scene_t"] == 32 and size["subject::scene_t"] == 48 and size["subject::material_t"] == 184
