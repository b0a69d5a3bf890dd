"""The Python examples of the integration guide (the repository's integration
notes at its root) run as written on the device."""
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GUIDE = "INTEGRATION.md"  # the file the examples are read from


def _python_blocks():
    text = open(os.path.join(ROOT, GUIDE)).read()
    return re.findall(r"```python\n(.*?)```", text, re.S)


def test_integration_examples_run(monkeypatch):
    blocks = _python_blocks()
    assert len(blocks) >= 2
    monkeypatch.chdir(ROOT)  # the raw ctypes example loads the library by its relative path
    for code in blocks:
        scope = {}
        exec(compile(code, GUIDE, "exec"), scope)
        if "x" in scope:
            x = scope["x"]
            assert x.shape == (3, 101) and np.isfinite(x).all()
