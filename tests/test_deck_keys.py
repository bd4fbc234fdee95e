"""The reference keys the XML surface accepts beyond the plain ones: each is either honoured or
refused with an error, never dropped (PoissonSolver.java:49-113, SolverModule.java:71-86,
MaxwellianSource.java:44-75, Source.java:68-71, Material.setInitValues, Material.java:580-608)."""
import math

import pytest

from starfish_amd import deck as D

DOMAIN = ('<domain type="xy"><mesh type="uniform" name="m"><origin>0,0</origin><spacing>0.01,0.01</spacing>'
          '<nodes>21,11</nodes><mesh-bc wall="left" type="dirichlet" value="0"/></mesh></domain>')
BOUNDARY = ('<boundaries><boundary name="inlet" type="virtual"><path>M 0,0.1 L 0,0</path></boundary>'
            '</boundaries>')
MATERIAL = ('<materials><material name="O+" type="kinetic_hip"><molwt>16</molwt><charge>1</charge>'
            '<spwt>1e6</spwt><init>nd_back=5, T=1500</init></material></materials>')


def _deck(tmp_path, solver="", source=""):
    (tmp_path / "sim.xml").write_text(
        "<simulation>%s%s%s<sources>%s</sources>%s<time><num_it>5</num_it><dt>1e-7</dt></time></simulation>"
        % (DOMAIN, BOUNDARY, MATERIAL, source, solver))
    return D.parse_deck(str(tmp_path / "sim.xml"))


def _source(body):
    return ('<boundary_source name="s" type="maxwellian"><material>O+</material><boundary>inlet</boundary>'
            '<v_drift>7000</v_drift><temperature>1000</temperature>%s</boundary_source>' % body)


def test_material_init_values(tmp_path):
    dk = _deck(tmp_path)
    assert dk.materials[0].nd_back == 5.0 and dk.materials[0].init_T == 1500.0


def test_source_rate_keys_current_and_den(tmp_path):
    mass = 16 * D.AMU
    dk = _deck(tmp_path, source=_source("<current>2e-3</current><v_az>250</v_az><start_it>3</start_it>"
                                        "<end_it>9</end_it>"))
    s = dk.sources[0]
    assert math.isclose(s.mdot, 2e-3 / D.QE * mass, rel_tol=1e-14)
    assert (s.v_az, s.start_it, s.stop_it) == (250.0, 3, 9)
    dk = _deck(tmp_path, source=_source("<den>1e14</den>"))
    v_mean = math.sqrt(8 * D.K_B * 1000 / (math.pi * mass))
    assert math.isclose(dk.sources[0].mdot, 1e14 * 0.1 * (7000 + v_mean / 4) * mass, rel_tol=1e-14)
    with pytest.raises(ValueError):
        _deck(tmp_path, source=_source("<mdot>1e-9</mdot><den>1e14</den>"))
    with pytest.raises(ValueError):
        _deck(tmp_path, source=_source(""))
    with pytest.raises(ValueError):
        _deck(tmp_path, source=_source("<mdot>1e-9</mdot><circuit_model>true</circuit_model>"))


def test_solver_keys(tmp_path):
    base = '<solver type="poisson_hip"><n0>1e10</n0><Te0>2</Te0><phi0>0</phi0>%s</solver>'
    dk = _deck(tmp_path, solver=base % "")
    assert dk.solver.skip == 1 and not dk.solver.initial_only and not dk.solver.nl_tol_set
    dk = _deck(tmp_path, solver=base % "<skip>3</skip><initial_only>true</initial_only><nl_tol>1e-6</nl_tol>")
    assert dk.solver.skip == 3 and dk.solver.initial_only and dk.solver.nl_tol_set and dk.solver.nl_tol == 1e-6
    for bad in ("<qn_switch>true</qn_switch>", "<n0_pos>0.1,0.1</n0_pos>", "<skip>0</skip>"):
        with pytest.raises(ValueError):
            _deck(tmp_path, solver=base % bad)
