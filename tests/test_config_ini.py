"""Run configuration files: the INI parser (IniConfig, modules/litho-io/src/io/ini_config.cpp)
and InversionConfig's mapping of it (ini_config_wiring.cpp:6-260), on hand-written text
and on the reference's two Forrestania configuration files (tests/golden/forrestania)."""
import inspect
import os

import numpy as np
import pytest

from lithoinv3d_amd import inversion, litho_io, synth

GOLD = os.path.join(os.path.dirname(__file__), "golden", "forrestania")

TEXT = """# comment
; another
orphan = before any section
[inversion]
max_iterations = 20
tolerance: 1e-5
flag = Yes
off = maybe
bad = abc
num = 5.7x
empty =
url = http://host:80/x = y
[broken
[gravity]
   density_min   =   1.5
dup = 1
dup = 2
"""


def test_ini_parsing_rules():
    c = litho_io.IniConfig(text=TEXT)
    e = c.entries()
    # lines before the first section are dropped; a header without ']' is skipped (the
    # keys after it stay in the current section); the first '=' splits, else ':'
    assert sorted(e) == ["gravity", "inversion"]
    assert e["inversion"]["tolerance"] == "1e-5"
    assert e["inversion"]["url"] == "http://host:80/x = y"
    assert e["inversion"]["empty"] == ""
    assert e["gravity"] == {"density_min": "1.5", "dup": "2"}
    assert c.get("nosuch", "key", "d") == "d" and c.get("inversion", "nosuch") is None
    assert c.get_int("inversion", "max_iterations", 0) == 20
    assert c.get_double("inversion", "tolerance", 0.0) == 1e-5
    # booleans: true / yes / 1 / on in any case, anything else false, empty -> default
    assert c.get_bool("inversion", "flag", False) is True
    assert c.get_bool("inversion", "off", True) is False
    assert c.get_bool("inversion", "empty", True) is True
    # numbers: the leading part (std::stod / std::stoi), default without one or if empty
    assert c.get_double("inversion", "bad", 3.5) == 3.5
    assert c.get_double("inversion", "num", 0.0) == 5.7
    assert c.get_int("inversion", "num", 0) == 5
    assert c.get_double("inversion", "empty", 2.0) == 2.0
    with pytest.raises(RuntimeError):
        litho_io.IniConfig(os.path.join(GOLD, "no_such_file.ini"))


def test_ini_save_and_reload(tmp_path):
    c = litho_io.IniConfig(text=TEXT)
    c.set("new", "k", "v w")
    c.save(tmp_path / "out.ini")
    assert litho_io.IniConfig(tmp_path / "out.ini").entries() == c.entries()
    text = (tmp_path / "out.ini").read_text().splitlines()
    assert text[0] == "[gravity]" and text[1] == "density_min = 1.5"


def test_forrestania_configuration_files():
    """The values of apps/forrestania/test_config.ini and test_config_real.ini by hand,
    and the reference's defaults for every key they leave out."""
    c = inversion.InversionConfig.from_ini(os.path.join(GOLD, "test_config.ini"))
    assert (c.max_iterations, c.tolerance, c.smoothness_weight, c.omega) == (20, 1e-5, 0.01, 2.0)
    assert (c.control_point_stride, c.lbfgs_history, c.vertex_freedom) == (5, 20, "z_only")
    assert (c.fd_step, c.armijo_c1, c.line_search_max_iter, c.solver) == (10.0, 1e-3, 30, "lbfgsb")
    assert c.enable_eigenvalue_scaling and c.enable_property_inversion
    assert (c.property_inversion_interval, c.property_inversion_max_iter) == (5, 10)
    assert (c.density_min, c.density_max, c.gravity_uncertainty) == (1.5, 5.0, 0.0)
    assert (c.magnetic_weight, c.inclination, c.declination, c.field_nT) == (0.5, -66.2, -0.1,
                                                                            58874.0)
    assert (c.susceptibility_min, c.susceptibility_max) == (0.0, 1.0)
    assert (c.remanence_min, c.remanence_max, c.remanence_mode) == (0.0, 50.0,
                                                                    "effective_susceptibility")
    assert c.enable_padding_group and c.padding_density_initial == 2.68
    assert (c.topography_mode, c.padding_cell_size) == ("none", 100.0)
    # not in the file: the reference's defaults
    assert (c.disable_line_search, c.enable_geometry_inversion, c.property_damping) == (
        False, True, 0.01)
    assert (c.gncg_cg_max_iter, c.gncg_cg_tolerance, c.mesh_boundary_mode) == (50, 1e-6, "free")
    assert (c.remanence_component_min, c.remanence_component_max) == (-10.0, 10.0)
    assert (c.enable_reference_model, c.lambda_ref, c.gradient_smoothing_weight) == (False, 0.1,
                                                                                     0.0)
    assert (c.enable_depth_bounds, c.depth_bound_margin, c.iteration_export_dir) == (False, 500.0,
                                                                                     "")
    r = inversion.InversionConfig.from_ini(os.path.join(GOLD, "test_config_real.ini"))
    assert (r.max_iterations, r.density_min, r.density_max, r.gravity_uncertainty) == (
        30, -0.5, 0.5, 0.015)
    assert (r.susceptibility_max, r.remanence_max) == (0.1, 10.0)


def test_enumerations_and_round_trip(tmp_path):
    c = inversion.InversionConfig.from_ini(litho_io.IniConfig(text="""[inversion]
vertex_freedom = diagonal
solver = gncg
[magnetic]
remanence_mode = vector
[topography]
mode = terrain_corrected
"""))
    # an unknown freedom is z_only, as fromIni's else branch makes it
    assert (c.vertex_freedom, c.solver, c.remanence_mode, c.topography_mode) == (
        "z_only", "gncg", "vector", "terrain_corrected")
    c.tolerance = 1.234567e-9  # beyond the six decimals of the reference's toIni
    c.iteration_export_dir = "out/iters"
    back = inversion.InversionConfig.from_ini(c.to_ini(tmp_path / "c.ini"))
    assert back == c
    assert inversion.InversionConfig.from_ini(tmp_path / "c.ini") == c
    assert inversion.InversionConfig.from_ini(inversion.InversionConfig().to_ini()) == \
        inversion.InversionConfig()


def test_settings_reach_model_problem_and_runner():
    c = inversion.InversionConfig.from_ini(os.path.join(GOLD, "test_config.ini"))
    # every keyword is one that Problem and Problem.run take
    assert set(c.problem_kwargs()) <= set(inspect.signature(inversion.Problem).parameters)
    assert set(c.run_kwargs()) <= set(inspect.signature(inversion.Problem.run).parameters)
    rk = c.run_kwargs()
    assert rk["density_bounds"] == (1.5, 5.0) and rk["property_interval"] == 5
    assert rk["solver"] == "lbfgsb" and rk["iteration_export_dir"] is None
    pk = c.problem_kwargs()
    assert (pk["inc"], pk["alpha_mag"], pk["smooth_weight"], pk["ref_weight"]) == (
        -66.2, 0.5, 0.01, 0.0)
    rng = np.random.default_rng(2)
    labels = rng.integers(0, 3, size=(8, 8, 4)).astype(np.int32)
    m = synth.label_grid_model(labels, (-100.0, -100.0, -50.0), (25.0, 25.0, 25.0),
                               [0.2, -0.1, 0.3], [0.01, 0.0, 0.02])
    M = inversion.Model(m)
    n = c.apply_to_model(M)
    M2 = inversion.Model(m)
    M2.fix_exterior_faces(inversion.Z_ONLY)
    M2.set_axis_scale(10.0)
    assert n == M2.set_cp_stride(5) > 0
    np.testing.assert_array_equal(M.params(), M2.params())
