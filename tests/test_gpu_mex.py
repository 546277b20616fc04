"""The MEX files on the GPU: the call sequence of the reference's Matlab class
(MultiGPUGridder_Matlab_Class.m:121-192 set-up, :233 forwardProject, :267
backProject, :320 getVol's CAS-volume copy) driven through the seven
adapters linked to libMultiGPUGridder.so (Matlab replaced by the test
stand-in of its MEX API), against the oracle."""
import os

import numpy as np
import pytest

from conftest import ROOT, assert_parity, rel_l2
from test_mex_dispatch import DOUBLE, INT32, LOGICAL, SINGLE, Matlab, build

pytestmark = pytest.mark.gpu


def test_matlab_call_sequence_through_the_mex_files(cuda, oracle, ker32, tmp_path):
    from multigpugridder_amd.synth import blob_volume, random_directions

    lib_dir = os.path.join(ROOT, "multigpugridder_amd")
    stub, libs = build(str(tmp_path), None, None,
                       extra=["-L", lib_dir, "-lMultiGPUGridder", f"-Wl,-rpath,{lib_dir}"])
    ml = Matlab(stub, libs)
    n, p = 32, 20
    m, v = 2 * n, 2 * n + 6
    print(f"MEX call sequence n={n} p={p}", flush=True)
    i32 = lambda *x: ml.arr(INT32, np.array(x, np.int32))
    sgl = lambda a: ml.arr(SINGLE, a)
    rc, err, h = ml("mexCreateGridder", i32(n), i32(p), sgl(np.array([2.0], np.float32)), i32(3),
                    ml.arr(DOUBLE, np.array([1.0])), i32(0), ml.arr(DOUBLE, np.array([1.0])),
                    ml.arr(LOGICAL, np.array([0], np.uint8)), nlhs=1)
    assert rc == 0, err
    H = lambda: ml.handle_arg(h)

    def setv(cmd, *vals):
        rc, err, _ = ml("mexSetVariables", ml.text(cmd), H(), *vals)
        assert rc == 0, (cmd, err)

    vol = blob_volume(n, 3)
    cas = oracle.vol_to_cas(vol, 2, 3).astype(np.float32)  # the class's host CASFromVol
    axes = random_directions(p, 21)
    images = np.zeros((p, n, n), np.float32)
    cas_out = np.zeros((v, v, v), np.float32)
    setv("SetCoordAxes", sgl(axes.reshape(-1)), i32(9 * p, 1))
    setv("SetVolume", sgl(vol), i32(n, n, n))
    setv("SetImages", sgl(images), i32(n, n, p))
    setv("SetGPUs", i32(0), i32(1))
    setv("SetKBTable", sgl(ker32), i32(1, ker32.size))
    setv("SetNumberStreamsFP", i32(10))
    setv("SetNumberStreamsBP", i32(4))
    setv("SetMaskRadius", sgl(np.array([m / 2 - 1], np.float32)))
    setv("SetNumAxes", i32(p))
    setv("SetCASVolume", sgl(cas), i32(v, v, v))
    setv("SetMaxAxesToAllocate", i32(1000))
    rc, err, _ = ml("mexMultiGPUForwardProject", H())
    assert rc == 0, err
    ref = oracle.cas_imgs_to_imgs(oracle.forward_project(cas, axes, ker32, m), n)
    assert_parity(images, ref, 1e-5, peak=1e-5)  # written in place into the borrowed array

    # back projection of those images into the (zeroed) CAS volume, copied out by GetVolume
    setv("SetCASVolume", sgl(cas_out), i32(v, v, v))
    rc, err, _ = ml("mexMultiGPUBackProject", H())
    assert rc == 0, err
    rc, err, _ = ml("mexMultiGPUGetVolume", H())
    assert rc == 0, err
    bp = oracle.back_project(oracle.imgs_to_cas_imgs(images, m).astype(np.float32), axes, ker32, v)
    assert rel_l2(cas_out, bp) < 1e-5

    rc, err, _ = ml("mexDeleteGridder", H())
    assert rc == 0, err
    assert ml.stub.stub_locks() == 0
