"""The controller's command line (controller/controller.cpp:35-96): its
positional arguments, defaults and codes, parsed on the host."""
from form_factors_amd import controller as ctl
from form_factors_amd import plugin as P


def test_defaults_and_positions():
    o, code = ctl.parse_args(["scene.obj", "task.task"])
    assert code == 0
    assert (o["result"], o["n_rays"], o["n_steps"], o["caster"], o["binary"]) == ("-", 1_000_000, 100, "cpu", False)
    o, _ = ctl.parse_args(["s", "t", "out.txt", "20000", "7", "naive_cuda", "1"])
    assert (o["result"], o["n_rays"], o["n_steps"], o["binary"]) == ("out.txt", 20000, 7, True)
    assert o["caster_type"] == P.RAY_CASTER_NAIVE_HIP
    assert ctl.parse_args(["s", "t", "-", "1", "1", "cuda"])[0]["caster_type"] == P.RAY_CASTER_GRID_HIP
    assert ctl.parse_args(["s", "t", "-", "1", "1", "cpu", "0"])[0]["binary"] is False


def test_atoi_semantics_and_errors():
    o, _ = ctl.parse_args(["s", "t", "-", "12abc", "x"])
    assert o["n_rays"] == 12 and o["n_steps"] == 0  # atoi reads leading digits, else 0
    assert ctl.parse_args(["only_scene"]) == (None, 1)  # usage
    assert ctl.parse_args(["s", "t", "-", "1", "1", "opencl"]) == (None, -2)  # -RAY_CASTER_NOT_SUPPORTED
    o, _ = ctl.parse_args(["s", "t", "--inline-materials"])
    assert o["inline_materials"] and o["result"] == "-"


def test_usage_exit_code(capsys):
    assert ctl.main(["x"]) == 1
    assert "usage" in capsys.readouterr().out
