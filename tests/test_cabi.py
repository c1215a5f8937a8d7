"""C ABI checks that need no GPU: the library loads and exports every entry
point declared in include/chidg_hip.h."""
import ctypes as C
import os
import re

import pytest

from chidg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(cdg_\w+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in sorted(decl) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == decl


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's size and offsets."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    checks = []
    structs = (("cdg_fgmres_params", _lib.FgmresParams), ("cdg_fgmres_stats", _lib.FgmresStats),
               ("cdg_request", _lib.Request), ("cdg_result", _lib.Result))
    for cname, py in structs:
        checks.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            checks.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src = tmp_path / "probe.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "chidg_hip.h"\nint main(void){'
                   + "".join(checks) + "return 0;}\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for cname, py in structs:
        assert int(got[cname]) == C.sizeof(py)
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_default_params_are_reference_defaults():
    lib = _lib.load()
    p = _lib.FgmresParams()
    lib.cdg_fgmres_default_params(C.byref(p))
    assert (p.nkrylov, p.tol, p.rtol, p.nmax, p.orthogonalization) == (2000, 1e-8, 1e-10, -1, 0)
    assert (p.inner_fgmres, p.inner_nkrylov, p.inner_tol, p.inner_rtol, p.inner_nmax) == (1, 100, 0.1, 0.1, 100)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        _lib.load()


def test_no_gpu_context_creation_errors_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.cdg_ctx_create(0, C.byref(h)) != 0
