"""The integration guide stays in step with the C ABI (CPU only): every
function include/vv1d.h declares is mapped there, and its Python binding
stubs are valid Python."""

import ast
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _read(name):
    with open(os.path.join(ROOT, name)) as fh:
        return fh.read()


def test_every_export_is_mapped():
    header = _read("include/vv1d.h")
    names = set(re.findall(r"^(?:int|int32_t|int64_t|const char \*)\s*\**\s*(vv_\w+)\(", header, re.M))
    assert len(names) > 50
    guide = _read("INTEGRATION.md")
    missing = sorted(n for n in names if n not in guide)
    assert not missing, missing


def test_binding_stubs_parse():
    blocks = re.findall(r"```python\n(.*?)```", _read("INTEGRATION.md"), re.S)
    assert blocks
    for code in blocks:
        ast.parse(code)
