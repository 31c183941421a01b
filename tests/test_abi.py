"""The C-ABI library loads and exports every function include/depsolver_amd.h declares."""
import ctypes
import os
import re

import depsolver_amd as dep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "depsolver_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|size_t|char\s*\*)\s*\**\s*(\w+)\s*\(",
                                 txt, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("dep_ctx_create", "dep_geom_create", "dep_factor", "dep_solve", "dep_gmres",
              "dep_post_mst", "electricFormFullA_tria6", "gaussBksb", "solverGMRES_el",
              "postProcess_tria6"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dep.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
