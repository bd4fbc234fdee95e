"""C-ABI checks that need no GPU: the library loads, exports every entry point declared in
include/starfish_hip.h, binds them all, and reports errors through status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from starfish_amd import _lib
from starfish_amd.plugin import Context, Geometry
from starfish_amd import deck as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "starfish_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_\w+)\s*\(", src)))


def test_every_declared_entry_point_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 30
    lib = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), "symbol %s not exported" % n
        assert n in _lib.SIGNATURES, "symbol %s has no ctypes binding" % n
    assert set(_lib.SIGNATURES) == set(names)


def test_version_and_null_arguments():
    lib = _lib.load()
    assert b"gfx950" in lib.sf_version()
    assert lib.sf_ctx_destroy(None) == _lib.SF_OK - 1    # SF_ERR_ARG
    assert lib.sf_ctx_create(0, None) < 0


def test_host_only_context_builds_geometry_and_refuses_device_work():
    ctx = Context(-1)
    try:
        deck = D.config_c1()
        g = Geometry(ctx, deck)
        assert g.node_type.shape == (deck.mesh.ni * deck.mesh.nj,)
        md = g.md
        rc = ctx.lib.sf_mesh_setup(ctx.h, C.byref(md), g.face_type.ctypes.data_as(C.POINTER(C.c_int)),
                                   None, g.node_type.ctypes.data_as(C.POINTER(C.c_int)), None,
                                   g.node_vol.ctypes.data_as(C.POINTER(C.c_double)), 0, None, None, None,
                                   g.node_seg_ptr.ctypes.data_as(C.POINTER(C.c_int)), None)
        assert rc == -2                                   # SF_ERR_HIP: a host-only context has no device
        assert b"hipSetDevice" in ctx.lib.sf_last_error(ctx.h)
        with pytest.raises(_lib.StarfishError):
            ctx.call("sf_push_deposit", 0)
    finally:
        ctx.close()


def test_geometry_rejects_degenerate_input():
    ctx = Context(-1)
    try:
        md = _lib.MeshDesc(0, 1, 5, (C.c_double * 2)(0, 0), (C.c_double * 2)(1, 1))
        ft = np.zeros(2 * 1 + 2 * 5, np.int32)
        g = C.c_void_p()
        rc = ctx.lib.sf_geometry_build(ctx.h, C.byref(md), ft.ctypes.data_as(C.POINTER(C.c_int)), None, 0, None,
                                       None, None, None, None, None, C.byref(g))
        assert rc == -1                                   # SF_ERR_ARG
    finally:
        ctx.close()


def test_deck_xml_round_trip_of_the_tutorial_surface(tmp_path):
    (tmp_path / "domain.xml").write_text(
        '<domain type="xy"><mesh type="uniform" name="mesh1"><origin>-0.15,0</origin>'
        '<spacing>5e-3, 5e-3</spacing><nodes>81, 41</nodes><mesh-bc wall="left" type="dirichlet" value="0" />'
        '<mesh-bc wall="bottom" type="symmetry"/></mesh></domain>')
    (tmp_path / "sim.xml").write_text(
        '<simulation><load>domain.xml</load><boundaries><boundary name="cyl" type="solid" value="-100">'
        '<material>SS</material><path>M 0.05, 0 L 0, -0.05 -0.05, 0 0, 0.05 0.05, 0</path></boundary>'
        '</boundaries><materials><material name="O+" type="kinetic_hip"><molwt>16</molwt><charge>1</charge>'
        '<spwt>1e3</spwt><init>nd_back=1e4</init></material></materials>'
        '<solver type="poisson_hip"><n0>1e10</n0><Te0>1.5</Te0><phi0>0</phi0><max_it>5e4</max_it></solver>'
        '<time><num_it>10</num_it><dt>5e-7</dt></time></simulation>')
    dk = D.parse_deck(str(tmp_path / "sim.xml"))
    assert (dk.mesh.ni, dk.mesh.nj) == (81, 41) and dk.mesh.mesh_bc["LEFT"] == ("DIRICHLET", 0.0)
    assert dk.boundaries[0].type == "DIRICHLET" and len(dk.boundaries[0].points) == 5
    assert dk.materials[0].nd_back == 1e4 and dk.materials[0].kind == "kinetic_hip"
    assert dk.solver.max_it == 50000 and dk.solver.Te0 == 1.5 and dk.num_it == 10 and dk.dt == 5e-7
