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


def _peak_stack(max_depth, children):
    """Peak size of one lane's ray stack under the Whitted kernel's protocol: pop a ray,
    push its children unless they would sit at max_depth (those return black, ray.rs:12-14,
    so the kernel never stores them)."""
    stack, peak = [0], 1
    while stack:
        d = stack.pop()
        if d + 1 < max_depth:
            stack.extend([d + 1] * children)
        peak = max(peak, len(stack))
    return peak


def test_ray_stack_bound_is_exact():
    """The capacity the host asks of the ray stack, 1 + (max_depth - 1) (children - 1)
    (whitted_impl in csrc/ft_backend.hip), is the protocol's peak: rough glass (7 children
    per hit) at the reference's default max_depth 8 needs 43 entries."""
    for max_depth in range(1, 9):
        for children in (1, 2, 4, 7):
            assert _peak_stack(max_depth, children) == 1 + (max_depth - 1) * (children - 1)
