"""The C-ABI libraries load and export every symbol their headers declare.

No compute calls here: the HIP library is only loaded (it initialises no
device until a context is created)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_device_library_exports_header():
    from flite3d_amd import _native
    lib = ctypes.CDLL(_native.device_path())
    names = declared("flite3d_amd.h", "flite_")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_prepro_library_exports_header():
    from flite3d_amd import _native
    lib = _native.prepro()
    names = declared("flite3d_prepro.h", "fp_")
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_structs_match_header_layout():
    from flite3d_amd.params import FliteLevelView, FliteParams
    text = open(os.path.join(ROOT, "include", "flite3d_amd.h")).read()
    body = re.search(r"typedef struct flite_params \{(.*?)\} flite_params;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(None, 1)[1].split(","):
            names.append(re.sub(r"\[.*", "", part).strip())
    assert [f[0] for f in FliteParams._fields_] == names
    body = re.search(r"typedef struct flite_level_view \{(.*?)\} flite_level_view;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    vnames = [re.sub(r"[\*\s]", "", d.split()[-1]) for d in body.split(";") if d.strip()]
    assert [f[0] for f in FliteLevelView._fields_] == vnames
