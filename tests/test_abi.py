"""The C-ABI library loads, exports every entry point include/ft_abi.h declares, and its
struct layouts match the Python mirror.  No compute call is made (no GPU needed)."""
import ctypes
import os
import re

from forgedthoughts_amd import abi, render

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "ft_abi.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(ft_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = render.library()
    names = declared_functions()
    assert "ft_scene_create" in names and "ft_render_whitted" in names
    for name in names:
        assert hasattr(lib, name), f"{name} declared in ft_abi.h but not exported"
    assert set(render.EXPORTS) <= set(names)


def test_struct_layout_matches_header():
    lib = render.library()
    for name, size in abi.STRUCT_SIZES.items():
        assert lib.ft_abi_struct_size(name.encode()) == size, name
    assert lib.ft_abi_struct_size(b"nope") == -1


def test_last_error_is_empty_string_initially():
    assert isinstance(render.library().ft_last_error(), bytes)
