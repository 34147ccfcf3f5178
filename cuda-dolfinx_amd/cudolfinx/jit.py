"""tabulate_tensor sources out of FFCx-generated C (python/cudolfinx/jit.py).

FFCx writes one C file per compiled form module: for every integral a
function

    void tabulate_tensor_integral_<id>(double* restrict A, const double* restrict w,
                                       const double* restrict c,
                                       const double* restrict coordinate_dofs,
                                       const int* restrict entity_local_index,
                                       const uint8_t* restrict quadrature_permutation,
                                       void* custom_data)
    { ... }

followed by its `ufcx_integral integral_<id>` record, and per form an array
`form_integrals_form_<name>[] = {&integral_<a>, &integral_<b>, ...}` giving the
order of the form's integrals.  The reference cuts the bodies out line by line
(function line starting with `void`, braces at column 0;
jit.py:15-59) and wraps each in its NVRTC header (jit.py:61-77, 90-114).
Here the C text is scanned with a small lexer that skips comments and string
and character literals, so the bodies are found by brace matching wherever
the braces sit; each body is re-wrapped in the engine's UFCx signature (the
six arguments above, without custom_data, which generated bodies of cell and
facet integrals do not read), ready for `mf_form_add_integral`, which makes
it a `__device__` function of the hipRTC kernels.
"""
from __future__ import annotations

import pathlib
import re
import typing

_FUNC = re.compile(r"\bvoid\s+tabulate_tensor_(integral_[A-Za-z0-9_]+)\s*\(")
_ORDER = re.compile(r"form_integrals_(form_[A-Za-z0-9_]+)\s*\[[^\]]*\]\s*=\s*\{")

SIGNATURE = ("void {name}({scalar}* restrict A, const {scalar}* restrict w, "
             "const {scalar}* restrict c, const {geom}* restrict coordinate_dofs, "
             "const int* restrict entity_local_index, "
             "const uint8_t* restrict quadrature_permutation)")


def _code_mask(text: str) -> bytearray:
    """1 for every character outside comments and string / character literals."""
    mask = bytearray(b"\x01") * len(text)
    i, n = 0, len(text)
    while i < n:
        ch = text[i]
        if ch == "/" and i + 1 < n and text[i + 1] == "/":
            j = text.find("\n", i)
            j = n if j < 0 else j
        elif ch == "/" and i + 1 < n and text[i + 1] == "*":
            j = text.find("*/", i + 2)
            j = n if j < 0 else j + 2
        elif ch in "\"'":
            j = i + 1
            while j < n and text[j] != ch:
                j += 2 if text[j] == "\\" else 1
            j = min(n, j + 1)
        else:
            i += 1
            continue
        mask[i:j] = b"\x00" * (j - i)
        i = j
    return mask


def _match(text: str, mask: bytearray, start: int, open_ch: str, close_ch: str) -> int:
    """Index of the bracket closing the one at `start` (code characters only)."""
    depth = 0
    for k in range(start, len(text)):
        if not mask[k]:
            continue
        if text[k] == open_ch:
            depth += 1
        elif text[k] == close_ch:
            depth -= 1
            if depth == 0:
                return k
    raise ValueError(f"unbalanced '{open_ch}' at offset {start}")


def _read(source) -> str:
    """C text of `source`: a DOLFINx-style form (its module's .c file, as
    jit.py:18-23 finds it), a path to a .c file, the C text itself, or an
    object holding it in a `text` attribute."""
    module = getattr(source, "module", None)
    if module is not None and getattr(module, "__file__", None):
        f = pathlib.Path(module.__file__)
        c = f.parent / (f.name.split(".")[0] + ".c")
        if not c.is_file():
            raise IOError(f"Could not find generated ffcx source file '{c}'!")
        return c.read_text()
    if isinstance(source, pathlib.Path) or (isinstance(source, str) and source.endswith(".c")
                                             and "\n" not in source):
        return pathlib.Path(source).read_text()
    if isinstance(source, str):
        return source
    text = getattr(source, "text", None)
    if isinstance(text, str):
        return text
    raise TypeError(f"expected a form, a .c path or C text, got '{type(source)}'")


def get_tabulate_tensor_sources(source, form_name: typing.Optional[str] = None):
    """[(integral id, body)] in order of appearance and, for the form's
    integrals in their `form_integrals_*` order, the index of each in that list
    (jit.py:15-59).  form_name selects one form of a multi-form file
    (`form_<name>`); by default the last array in the file, as the reference."""
    text = _read(source)
    mask = _code_mask(text)
    bodies = []
    for m in _FUNC.finditer(text):
        if not mask[m.start()]:
            continue                                     # inside a comment or string
        close = _match(text, mask, m.end() - 1, "(", ")")
        brace = close + 1
        while brace < len(text) and (text[brace].isspace() or not mask[brace]):
            brace += 1
        if brace >= len(text) or text[brace] != "{":
            continue                                     # a declaration, not a definition
        end = _match(text, mask, brace, "{", "}")
        bodies.append((m.group(1).replace("integral_", "", 1), text[brace + 1:end]))
    orders = {}
    for m in _ORDER.finditer(text):
        if not mask[m.start()]:
            continue
        end = _match(text, mask, m.end() - 1, "{", "}")
        items = [t.strip() for t in text[m.end():end].split(",") if t.strip()]
        orders[m.group(1)] = [t.lstrip("&").strip().replace("integral_", "", 1) for t in items]
    if not bodies:
        raise ValueError("no tabulate_tensor_integral_* definitions found")
    if form_name is not None:
        key = form_name if form_name.startswith("form_") else "form_" + form_name
        if key not in orders:
            raise KeyError(f"no form_integrals_{key} array in the source")
        ids = orders[key]
    elif orders:
        ids = list(orders.values())[-1]
    else:
        ids = [b[0] for b in bodies]
    position = {bid: i for i, (bid, _) in enumerate(bodies)}
    missing = [i for i in ids if i not in position]
    if missing:
        raise ValueError(f"integrals {missing} have no tabulate_tensor definition")
    return bodies, [position[i] for i in ids]


def get_wrapped_tabulate_tensors(source, form_name: typing.Optional[str] = None,
                                 scalar_type: str = "double"):
    """[(function name, C function in the engine's UFCx signature)] and the
    integral order (jit.py:90-114).  Only double precision is supported: the
    engine's geometry and coefficients are FP64 (the reference binds only
    <double, double> too, fem.cpp:399-402)."""
    if scalar_type != "double":
        raise NotImplementedError(f"scalar type '{scalar_type}' is not supported")
    bodies, order = get_tabulate_tensor_sources(source, form_name)
    out = []
    for bid, body in bodies:
        name = "tabulate_tensor_integral_" + bid
        head = SIGNATURE.format(name=name, scalar=scalar_type, geom=scalar_type)
        out.append((name, head + "\n{" + body + "}\n"))
    return out, order
