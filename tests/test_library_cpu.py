"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, computes the host math like the oracle, and reports errors
as status codes.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, gpu_available

import multigpugridder_amd as mgg
from multigpugridder_amd import _lib

HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def header_functions():
    text = "\n".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"defined", "if", "sizeof"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    declared = header_functions()
    assert len(declared) > 60
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding declares signatures for all of them
    assert set(declared) <= set(_lib.EXPORTED)


def test_version_and_error_string():
    assert b"gfx950" in _lib.lib().mgg_version()
    assert isinstance(_lib.last_error(), str)


def test_missing_library_fails_loudly():
    """Without the built library every entry point raises: there is no CPU
    fallback behind the gridder class, the operators or the server."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from multigpugridder_amd import _lib\n"
        "_lib.LIB_PATH = '/nonexistent/libMultiGPUGridder.so'\n"
        "from multigpugridder_amd import MultiGPUGridder, ops\n"
        "import numpy as np\n"
        "for f in (lambda: MultiGPUGridder(16, 4, 2), lambda: _lib.call('mgg_version'),\n"
        "          lambda: ops.kb_table(device='cpu')):\n"
        "    try:\n"
        "        f()\n"
        "    except _lib.LibraryNotBuilt as e:\n"
        "        assert 'no CPU fallback' in str(e)\n"
        "    else:\n"
        "        raise SystemExit('no error')\n"
        "print('refused')\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "refused" in r.stdout, r.stdout + r.stderr


def test_kb_table_matches_oracle(oracle):
    ours = mgg.kb_table(2, 501)
    ref = oracle.kb_table(2, 501).astype(np.float32)
    assert np.max(np.abs(ours.view(np.int32) - ref.view(np.int32))) <= 1


@pytest.mark.parametrize("m", [32, 128, 256, 384, 512])
def test_precomp_matches_oracle(oracle, m):
    np.testing.assert_allclose(mgg.precomp(m), oracle.precomp(m), rtol=2e-7)


def test_uniform_axes_match_oracle(oracle):
    np.testing.assert_allclose(mgg.uniform_axes(10, 50), oracle.uniform_axes(10, 50), atol=2e-7)
    a = mgg.uniform_axes(100, 100)
    assert a.shape == (10000, 9)


def test_host_math_rejects_bad_arguments():
    out = np.zeros(501, np.float32)
    with pytest.raises(_lib.MggError) as e:
        _lib.call("mgg_kb_table", out.ctypes.data, 501, 9)
    assert e.value.code == _lib.MGG_ERR_INVALID_ARGUMENT
    with pytest.raises(_lib.MggError):
        _lib.call("mgg_precomp", out.ctypes.data, 7, 2)  # odd size
    with pytest.raises(_lib.MggError):
        _lib.call("mgg_uniform_axes", None, 1, 1)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_gpu_reports_device_error():
    h = ctypes.c_void_p()
    rc = _lib.lib().mgg_create(64, 10, 2.0, 3, 1, None, 1, 0, ctypes.byref(h))
    assert rc == _lib.MGG_ERR_DEVICE
    assert "GPU" in _lib.last_error()
    assert _lib.lib().Gridder_new() is None


def test_create_rejects_bad_geometry():
    h = ctypes.c_void_p()
    lib = _lib.lib()
    assert lib.mgg_create(0, 10, 2.0, 3, 1, None, 1, 0, ctypes.byref(h)) == _lib.MGG_ERR_INVALID_ARGUMENT
    assert lib.mgg_create(64, 10, 0.0, 3, 1, None, 1, 0, ctypes.byref(h)) == _lib.MGG_ERR_INVALID_ARGUMENT
    assert lib.mgg_create(64, 10, 2.0, -1, 1, None, 1, 0, ctypes.byref(h)) == _lib.MGG_ERR_INVALID_ARGUMENT
    assert lib.mgg_create(64, 10, 2.0, 3, 2, None, 1, 0, ctypes.byref(h)) == _lib.MGG_ERR_UNSUPPORTED
    assert "one GPU" in _lib.last_error()


def test_setters_reject_null_handle_and_bad_values():
    lib = _lib.lib()
    assert lib.mgg_set_num_axes(None, 3) == _lib.MGG_ERR_INVALID_ARGUMENT
    assert lib.mgg_forward_project(None) == _lib.MGG_ERR_INVALID_ARGUMENT
    assert lib.mgg_destroy(None) == _lib.MGG_OK


def test_device_ops_validate_before_launch():
    lib = _lib.lib()
    # null pointers and odd sizes are rejected on the host, before any launch
    rc = lib.mgg_op_forward_project(None, 0, 38, None, 4, None, 32, 15.0, None, 501, 2.0, None)
    assert rc == _lib.MGG_ERR_INVALID_ARGUMENT
    rc = lib.mgg_op_forward_project_checked(None, 0, 38, None, 4, None, 32, 15.0, None, 501, 2.0, None, None)
    assert rc == _lib.MGG_ERR_INVALID_ARGUMENT
    rc = lib.mgg_op_check_volume(None, 0, 38, None, 501, None, None)
    assert rc == _lib.MGG_ERR_INVALID_ARGUMENT
    rc = lib.mgg_op_check_images(None, 4, 32, None, 501, None, None)
    assert rc == _lib.MGG_ERR_INVALID_ARGUMENT
    rc = lib.mgg_op_back_project_adjoint_checked(None, 38, None, 4, None, 32, 15.0, None, 501, 2.0, None, None)
    assert rc == _lib.MGG_ERR_INVALID_ARGUMENT
    rc = lib.mgg_op_cas_images_to_images(None, None, 4, 33, 16, None)
    assert rc == _lib.MGG_ERR_INVALID_ARGUMENT
    rc = lib.mgg_filter_divide_scalar(None, ctypes.c_float(0.0), 10, None)
    assert rc == _lib.MGG_ERR_INVALID_ARGUMENT
    assert "divide by zero" in _lib.last_error()


def test_python_host_fft_path_matches_oracle(oracle):
    """The RunFFTOnGPU=0 host path (utils/imgsFromCASImgs.m, CASImgsFromImgs.m)."""
    from multigpugridder_amd.hostmath import cas_imgs_from_imgs, imgs_from_cas_imgs

    rng = np.random.default_rng(0)
    c = rng.standard_normal((3, 32, 32)).astype(np.float32)
    np.testing.assert_allclose(imgs_from_cas_imgs(c, 16), oracle.cas_imgs_to_imgs(c, 16), atol=1e-6)
    x = rng.standard_normal((3, 16, 16)).astype(np.float32)
    np.testing.assert_allclose(cas_imgs_from_imgs(x, 32), oracle.imgs_to_cas_imgs(x, 32), rtol=1e-4, atol=1e-4)


def test_timed_cpu_path_matches_oracle(oracle, ker32):
    """The FP32 OpenMP path timed as the CPU baseline agrees with the FP64
    oracle to FP32 accumulation error."""
    from conftest import rel_l2
    from multigpugridder_amd.synth import blob_volume, random_directions

    cas = oracle.vol_to_cas(blob_volume(16, 1), 2, 3).astype(np.float32)
    axes = random_directions(6, 1)
    assert rel_l2(oracle.cpu_forward_project_f32(cas, axes, ker32, 32, nthreads=2),
                  oracle.forward_project(cas, axes, ker32, 32)) < 1e-6
    y = np.random.default_rng(1).standard_normal((6, 32, 32)).astype(np.float32)
    assert rel_l2(oracle.cpu_back_project_f32(y, axes, ker32, 38, nthreads=2),
                  oracle.back_project(y, axes, ker32, 38)) < 1e-6
