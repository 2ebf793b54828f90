"""The C example of the integration notes against include/hmv.h: the
fragment, wrapped in a function whose parameters give its free names their
types, must compile cleanly with gcc (every call matches its prototype:
argument counts, pointer types, no implicit declarations).  CPU only."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

WRAP = """#include <stdio.h>
#include "hmv.h"
void example(int nd, int nlf, const void *leaves, const double *d_x, double *d_y, void *stream,
             const double *d_b, double *d_u, hmv_arena a_bf16, const double *d_X, double *d_Y,
             const double *d_B, double *d_U, hmv_column_info *cols, int ninst, const int *nds,
             const int *nlfs, const void *const *leaf_ptrs, hmv_column_info *inst,
             double *inst_seconds) {
%s
}
"""


def _example():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## A C caller using the engine API directly"):]
    m = re.search(r"```c\n(.*?)```", sec, re.S)
    assert m, "no C block in the section"
    return m.group(1)


def test_integration_c_example_matches_the_header(tmp_path):
    if not subprocess.run(["which", "gcc"], capture_output=True).returncode == 0:
        pytest.skip("gcc not found")
    src = tmp_path / "example.c"
    src.write_text(WRAP % _example())
    p = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", "-Werror=implicit-function-declaration",
                        "-Werror=incompatible-pointer-types", "-Werror=int-conversion",
                        "-Wno-unused-variable", "-Wno-unused-but-set-variable",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    assert "warning" not in p.stderr, p.stderr
