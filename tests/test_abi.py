"""The C-ABI library loads (no GPU needed) and exports every function that
include/infraga_gpu.h declares; struct layouts seen from Python match."""
import ctypes as C
import os
import re
import subprocess

from infraga_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "infraga_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ga_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_expected_api():
    names = declared()
    for n in engine.EXPORTED:
        assert n in names
    assert len(names) == len(engine.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for n in declared():
        assert n in exported, n
        getattr(lib, n)


def test_struct_layouts_match_the_c_header(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's size and
    field offsets for include/infraga_gpu.h."""
    hdr = os.path.join(ROOT, "include", "infraga_gpu.h")
    structs = {"ga_prop_params": engine.PropParams, "ga_arrival": engine.Arrival, "ga_run_stats": engine.RunStats,
               "ga_eig_params": engine.EigParams, "ga_request": engine.Request, "ga_result": engine.Result,
               "ga_path_pt": engine.PathPt}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % hdr, "int main(void) {"]
    for cname, cls in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-o", str(exe), str(src)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


def test_defaults_follow_reference_driver():
    p = engine.prop_params(engine.GA_GEOM_SPH)
    assert (p.s_max, p.ds_min, p.ds_max, p.rng_max, p.time_max) == (1000.0, 0.001, 0.05, 1000.0, 10.0)
    assert (p.freq, p.abs_coeff, p.turn_ht_min, p.bounces, p.r_earth) == (0.1, 1.0, 0.2, 10, 6370.0)
    q = engine.prop_params(engine.GA_GEOM_3D)
    assert q.rng_max == 2000.0 and list(q.bounds) == [-2000.0, 2000.0, -2000.0, 2000.0]


def test_no_device_reports_error():
    """Without a GPU (this container) context creation fails cleanly."""
    lib = engine.load_library()
    h = C.c_void_p()
    rc = lib.ga_ctx_create(C.byref(h), 99)
    assert rc == engine.GA_E_NODEV and not h.value
