"""The C-ABI library loads and exports every entry point include/sbs/sbs.h declares."""
import ctypes
import re

from sbs_amd import capi


def declared_functions():
    text = open(capi.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sbs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("sbs_pd_create", "sbs_pd_update", "sbs_pd_set_global_solver", "sbs_pd_solve",
                     "sbs_pd_reset", "sbs_pd_set_perf", "sbs_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_functions()) == set(capi.SIGNATURES)


def test_errors_are_reported_not_raised():
    L = capi.lib()
    # invalid arguments are rejected without touching a GPU
    out = ctypes.c_void_p()
    assert L.sbs_pd_create(16, 128, None, ctypes.byref(out)) == -1
    assert b"null" in L.sbs_last_error()
    d = capi.SbsSolverData()
    assert L.sbs_pd_create(16, 128, ctypes.byref(d), ctypes.byref(out)) == -1
    assert b"precision" in L.sbs_last_error()
    assert L.sbs_pd_set_global_solver(None, 0) == -1
    assert L.sbs_version() >= 1


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct of the header has the same size and field offsets in C (gcc)
    as in the ctypes binding."""
    import subprocess
    structs = {"SbsSolverData": capi.SbsSolverData, "SbsSolverParams": capi.SbsSolverParams,
               "SbsPdStats": capi.SbsPdStats, "SbsFixedBody": capi.SbsFixedBody, "SbsIpcData": capi.SbsIpcData,
               "SbsIpcParams": capi.SbsIpcParams, "SbsIpcStats": capi.SbsIpcStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{capi.HEADER_PATH}"', "int main(void) {"]
    expect = []
    for name, cls in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({name}));')
        expect.append(ctypes.sizeof(cls))
        for field, _ in cls._fields_:
            cfield = "lambda" if field == "lambda_" else field
            lines.append(f'printf("%zu\\n", offsetof({name}, {cfield}));')
            expect.append(getattr(cls, field).offset)
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == expect
