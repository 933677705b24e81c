"""CPU tests of the C ABI libraries: every symbol declared in include/*.h is
exported, struct layouts match the reference, INI parsing and host-side
validation work without a GPU (no compute call is made here)."""
import ctypes
import os
import re
import subprocess

import pytest

from rfba_amd import srrf, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("sizeof",)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    return set(l.split()[-1] for l in out.splitlines() if l.strip())


@pytest.mark.parametrize("header,lib", [("rfba_srrf.h", "rfba_srrf"), ("rfba_3b.h", "rfba_3b"),
                                        ("rfba_prep.h", "rfba_prep")])
def test_all_declared_symbols_exported(header, lib):
    if not os.path.exists(os.path.join(ROOT, "include", header)):
        pytest.skip("header not present")
    names = declared_functions(header)
    assert len(names) > 3
    syms = exported(_lib.lib_path(lib))
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    ctypes.CDLL(_lib.lib_path(lib))  # loads without a GPU


def test_srrfconf_layout_matches_reference():
    # kernel.cuh:3-24: one float followed by 16 ints
    assert ctypes.sizeof(srrf.SRRFConf) == 68
    assert srrf.SRRFConf.RingRadius.offset == 0
    assert srrf.SRRFConf.blockSizeZ.offset == 64


def test_defaults_match_reference():
    c = srrf.default_conf()
    # GradientCalc.cpp:19-41
    assert c.RingRadius == 0.5 and c.Magnification == 5 and c.AxesinRing == 6 and c.MainRange == 1
    assert (c.XRange, c.YRange, c.ZRange) == (0, 0, 0)
    assert (c.Method, c.ITC, c.TRACOrder) == (1, 1, 2)
    assert (c.RemoveCon, c.Renormalize, c.GradientSmooth, c.IntensityW) == (0, 0, 0, 1)
    assert (c.blockSizeX, c.blockSizeY, c.blockSizeZ) == (16, 16, 16)


SHIPPED_INI = """[SRRF]
RingRadius=0.5
Magnification=5
AxesinRing=6
MainRange=1
XRange=1
YRange=1
ZRange=1
[Temproal]
#0:TRM
#1:TRA
#2:TRPPM
#3:TRAC
Method=1
#If use TRAC
Integrate_Temporal_Correlations=1
TRAC_ORDER=2

[Radiality]
Remove_Positivity_Constraint=0
Renormalize=0
Gradient_Smoothing=1

[Weighting]
Intensity_Weighting=1
"""


def test_ini_shipped_file(tmp_path):
    # the shipped x64/Release/SRRF.ini overrides X/Y/ZRange and smoothing
    p = tmp_path / "SRRF.ini"
    p.write_text(SHIPPED_INI)
    c = srrf.load_ini(str(p))
    assert (c.XRange, c.YRange, c.ZRange, c.GradientSmooth) == (1, 1, 1, 1)
    assert c.Magnification == 5 and c.Method == 1 and c.ITC == 1 and c.TRACOrder == 2
    assert c.blockSizeX == 16  # no [Split] section -> default


def test_ini_missing_file_defaults(tmp_path):
    c = srrf.load_ini(str(tmp_path / "absent.ini"))
    d = srrf.default_conf()
    assert c.as_dict() == d.as_dict()


def test_ini_case_and_comments(tmp_path, monkeypatch):
    p = tmp_path / "SRRF.ini"
    p.write_text("; comment\n[srrf]\nMAGNIFICATION = 3\nringradius=0.75\n[Split]\nBlockSizeX=24\n"
                 "[SRRF]\nMagnification=9\n")
    c = srrf.load_ini(str(p))
    assert c.Magnification == 3 and c.RingRadius == 0.75 and c.blockSizeX == 24
    monkeypatch.chdir(tmp_path)
    assert srrf.SRRFGetMag() == 3


def test_invalid_arguments_rejected_on_host():
    import numpy as np
    c = srrf.default_conf().replace(Magnification=0)
    with pytest.raises(srrf.SRRFError):
        srrf.srrf_stack(np.zeros((1, 4, 4, 4), np.uint16), c)
    c = srrf.default_conf().replace(Method=7)
    with pytest.raises(srrf.SRRFError):
        srrf.srrf_stack(np.zeros((1, 4, 4, 4), np.uint16), c)
    with pytest.raises(srrf.SRRFError):
        srrf.directions(5, 0.5)
    with pytest.raises(srrf.SRRFError):
        srrf.ring(0, 0.5)


def test_cxx_linkage_entry_points(tmp_path):
    # GradientCalc.h:4-5 declares both entry points with C++ linkage, and the
    # callers import them so (multispot5_place_methods.cc:48-49,
    # multispot5.cc:77): the library exports the mangled names too, and a
    # caller compiled from that declaration (dllimport removed) links and runs.
    syms = exported(_lib.lib_path("rfba_srrf"))
    assert {"_Z16SRRFGradientCalcPfPiS_", "_Z10SRRFGetMagv", "SRRFGradientCalc", "SRRFGetMag",
            "SRRFGradientCalcGPU"} <= syms
    src = tmp_path / "caller.cpp"
    src.write_text("#include <cstdio>\n"
                   "int SRRFGradientCalc(float * inImage, int * imageDim, float * outImage);\n"
                   "int SRRFGetMag();\n"
                   "int main() { std::printf(\"%d\\n\", SRRFGetMag());"
                   " return SRRFGradientCalc == nullptr; }\n")
    exe = tmp_path / "caller"
    libdir = os.path.dirname(_lib.lib_path("rfba_srrf"))
    r = subprocess.run(["g++", str(src), "-o", str(exe), "-L" + libdir, "-l:librfba_srrf.so",
                        "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    (tmp_path / "SRRF.ini").write_text("[SRRF]\nMagnification=3\n")
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=str(tmp_path))
    assert r.returncode == 0 and r.stdout.strip() == "3", (r.returncode, r.stdout, r.stderr)
