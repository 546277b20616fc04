"""The public headers are plain C: a C99 program that uses every declared type
and takes the address of every declared function compiles with -Wall -Werror
and links against libMultiGPUGridder.so (nothing is run, no GPU needed)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT
from test_library_cpu import header_functions

LIB_DIR = os.path.join(ROOT, "multigpugridder_amd")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_header_compiles_and_links_as_c99(tmp_path):
    names = header_functions()
    src = ['#include "mgg.h"', '#include "mgg_mex.h"', "#include <stdio.h>", "int main(void) {",
           "    mgg_mex_arg marg = {0}; (void)marg;",
           "    mgg_server_config cfg = {0}; mgg_request req = {0}; mgg_request_info info = {0};",
           "    mgg_server_stats st = {0}; mgg_gridder *g = NULL; mgg_server *s = NULL;",
           "    (void)cfg; (void)req; (void)info; (void)st; (void)g; (void)s;",
           "    void *fns[] = {"]
    src += [f"        (void *){n}," for n in names]
    src += ["    };", '    printf("%d\\n", (int)(sizeof fns / sizeof fns[0]));',
            "    return MGG_OK + MGG_REQ_FORWARD + MGG_DTYPE_F32;", "}"]
    c = tmp_path / "uses_header.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "uses_header"
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-Wno-pedantic", f"-I{os.path.join(ROOT, 'include')}",
           str(c), f"-L{LIB_DIR}", "-lMultiGPUGridder", f"-Wl,-rpath,{LIB_DIR}", "-Wl,--allow-shlib-undefined",
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()
