"""The C-ABI library loads, exports every symbol include/pic_amd.h declares, and
its host-only entry points behave (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pic_amd.h")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(pic_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pic_sim_step" in names and "pic_solve_multigrid" in names and len(names) >= 40


def test_library_exports_every_declared_symbol(pic):
    lib = ctypes.CDLL(pic.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header(pic):
    assert sorted(pic.SIGNATURES) == declared_functions()


STRUCTS = [("pic_mg_params", "MGParams"), ("pic_mg_stats", "MGStats"), ("pic_xs_process", "XsProcess"),
           ("pic_sim_params", "SimParams"), ("pic_sim_report", "SimReport"), ("pic_mcc_info", "MccInfo"),
           ("pic_diagnostics", "Diagnostics"), ("pic_tune_report", "TuneReport"), ("pic_job_desc", "JobDesc"),
           ("pic_job_result", "JobResult")]


def test_struct_layouts_match_header(pic, tmp_path):
    """Every structure of the ABI has the same size and field offsets in the
    header (compiled by gcc) as in its Python mirror."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pic_amd.h"', "int main(void) {"]
    want = []
    for cname, pyname in STRUCTS:
        cls = getattr(pic, pyname)
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        want.append(ctypes.sizeof(cls))
        for f in cls._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, f[0]))
            want.append(getattr(cls, f[0]).offset)
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want


def test_read_cross_section_matches_file(pic):
    path = os.path.join(GOLDEN, "csf_Ar.txt")
    e, s = pic.read_cross_section(path)
    ref = np.loadtxt(path)
    assert len(e) == len(ref) == 282
    assert np.array_equal(e, ref[:, 0]) and np.array_equal(s, ref[:, 1])
    assert e[0] == 0.0 and s[0] == 6.2984e-20          # csf/Ar.txt:1
    e2, s2 = pic.read_cross_section(os.path.join(GOLDEN, "csf_Ar+.txt"))
    assert e2[0] == 15.76 and s2[0] == 0.0             # ionization threshold row


def test_read_cross_section_rejects_latex_format(pic):
    # csf/Ar_conv.txt writes numbers as 6.2984*10^{-20}; the parser keeps no row
    e, s = pic.read_cross_section(os.path.join(GOLDEN, "csf_Ar_conv_head.txt"))
    assert len(e) == 0


def test_read_cross_section_skips_comments(pic, tmp_path):
    p = tmp_path / "xs.txt"
    p.write_text("# header\n\n1.0 2e-20\n# mid\n3.0\t4e-20\n")
    e, s = pic.read_cross_section(str(p))
    assert list(e) == [1.0, 3.0] and list(s) == [2e-20, 4e-20]


def test_missing_file_is_an_error(pic):
    with pytest.raises(pic.PicError):
        pic.read_cross_section("/nonexistent/xs.txt")


def test_mg_presets(pic):
    ref = pic.mg_params(pic.PRESET_REFERENCE)
    assert (ref.pre, ref.post, ref.coarse_sweeps, ref.omega, ref.tol, ref.max_cycles, ref.mode) == \
        (3, 3, 20, 1.25, 1e-5, 50, pic.MG_FP64)
    st = pic.mg_params(pic.PRESET_STEP)
    assert (st.pre, st.post, st.omega, st.tol, st.mode) == (2, 2, 1.0, 1e-8, pic.MG_MIXED)
    sor = pic.mg_params(pic.PRESET_STEP_SOR)
    assert (sor.pre, sor.post, sor.omega, sor.tol, sor.mode, sor.inner) == (2, 2, 1.2, 1e-8, pic.MG_MIXED, 2)
    poi = pic.mg_params(pic.PRESET_POISSON)
    assert (poi.omega, poi.inner) == (1.2, 2)
    assert ref.inner == 1 and st.inner == 2
    assert ref.coarse_direct == 0 and st.coarse_direct == sor.coarse_direct == poi.coarse_direct == 2
    with pytest.raises(pic.PicError):
        pic.mg_params(7)


def test_sim_default_params(pic):
    p = pic.sim_params()
    assert p.bin_w == 16 and p.halo == 2 and p.eps0 == 8.8541878128e-12
    assert p.mg.omega == 1.2 and p.mg.tol == 1e-8


def test_last_error_reports_message(pic):
    lib = pic.load()
    rc = lib.pic_xs_set_table(None, 1, 1.0)
    assert rc == -1
    assert b"bad table" in lib.pic_last_error()


def test_bench_refuses_more_than_one_gpu():
    """The benchmark runs on one GPU; asked for more it exits with status 2 and
    says why (before touching a device)."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr and not out.stdout.strip()


def test_missing_library_fails_loudly(pic, tmp_path):
    """Without the built HIP library the package raises; it has no CPU path."""
    with pytest.raises(RuntimeError, match="libpic_amd.so not found"):
        pic.load(str(tmp_path / "libpic_amd.so"))
