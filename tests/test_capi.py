"""The C ABI on the CPU: the library loads and exports every declared symbol.

No compute call is made here (there is no GPU in this suite).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(skinny_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_is_built():
    import skinny_matmul_amd as sk
    assert sk.available(), "run python -m skinny_matmul_amd.build"
    sk.lib()


def test_exports_every_declared_symbol():
    import skinny_matmul_amd as sk
    out = subprocess.run(["nm", "-D", "--defined-only", sk.library_path()],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    want = declared_functions()
    assert want, "no declarations found"
    missing = sorted(want - exported)
    assert not missing, missing


def test_python_binding_covers_header():
    import skinny_matmul_amd as sk
    missing = sorted(declared_functions() - set(sk.EXPORTED))
    assert not missing, missing


def test_status_strings():
    import skinny_matmul_amd as sk
    L = sk.lib()
    assert L.skinny_status_string(0) == b"ok"
    assert L.skinny_status_string(1) == b"unsupported"


def test_no_oracle_in_product():
    """The product path never imports or links the oracle."""
    pkg = os.path.join(ROOT, "skinny_matmul_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "orc_" not in text, f
    import skinny_matmul_amd as sk
    out = subprocess.run(["ldd", sk.library_path()], capture_output=True, text=True).stdout
    assert "oracle" not in out


def test_fill_rejects_mismatched_tensors():
    """The fill routines write n elements of their own type from an address;
    the binding refuses a tensor of another element size or fewer than n
    elements before anything reaches the device (a float32 tensor passed to
    the double fill would be written past its end)."""
    torch = pytest.importorskip("torch")
    import skinny_matmul_amd as sk
    f32 = torch.empty(16, dtype=torch.float32)
    f64 = torch.empty(16, dtype=torch.float64)
    with pytest.raises(sk.SkinnyError, match="4 bytes"):
        sk.fill_uniform(f32, 16, 1, 0)
    with pytest.raises(sk.SkinnyError, match="8 bytes"):
        sk.fill_uniform_f(f64, 16, 1, 0)
    with pytest.raises(sk.SkinnyError, match="into a tensor of 16"):
        sk.fill_int(f64, 17, 1, 0)
    # an expanded view: 16 elements over 8 bytes of storage
    with pytest.raises(sk.SkinnyError, match="not contiguous"):
        sk.fill_uniform(torch.empty(1, dtype=torch.float64).expand(16), 16, 1, 0)
    # a strided column view: its neighbours would be overwritten
    with pytest.raises(sk.SkinnyError, match="not contiguous"):
        sk.fill_uniform(torch.empty(8, 4, dtype=torch.float64)[:, 0], 8, 1, 0)
    # a contiguous view near the end of its storage
    tail = torch.empty(16, dtype=torch.float64)[12:]
    with pytest.raises(sk.SkinnyError, match="into a tensor of 4"):
        sk.fill_uniform(tail, 5, 1, 0)
