"""CPU checks of the C-ABI library: it loads and exports every entry point
declared in include/cup2d.h (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cup2d.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(cup2d_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "cup2d_create" in names and "cup2d_simulate" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    import cubismup_2d_amd as cup
    lib = ctypes.CDLL(cup.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(cup.EXPORTS)


def test_no_cpu_fallback_in_package():
    """The product module only binds the HIP library; it never imports the oracle."""
    src = open(os.path.join(ROOT, "cubismup_2d_amd", "__init__.py")).read()
    assert "import oracle" not in src and "from oracle" not in src
