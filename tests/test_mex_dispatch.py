"""The MEX layer (include/mgg_mex.h, multigpugridder_amd/mex) driven as Matlab
drives the reference's MEX files (src/Mex/*.cpp), on the CPU: the seven
adapters are compiled against a stand-in for Matlab's MEX API (tests/mex) and
linked to the dispatch layer over a recording double of the gridder C ABI, so
every check (argument counts and classes, the uint64 handle convention of
Mex/mexFunctionWrapper.h:29-82, the command dispatch of
Mex/mexSetVariables.cpp:27-97, status -> error text) runs here, and each
command is checked to reach the C ABI call it stands for with its arguments."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

MEX_DIR = os.path.join(ROOT, "multigpugridder_amd", "mex")
STUB_DIR = os.path.join(ROOT, "tests", "mex")
ENTRIES = ["mexCreateGridder", "mexSetVariables", "mexMultiGPUForwardProject", "mexMultiGPUBackProject",
           "mexMultiGPUGetVolume", "mexMultiGPUReconstructVolume", "mexDeleteGridder"]
# mxClassID values of the stand-in header (Matlab's order)
LOGICAL, CHAR, DOUBLE, SINGLE, INT32, UINT64 = 3, 4, 6, 7, 12, 15


def build(tmp, backend_srcs, backend_name, extra=()):
    """Compiles the MEX stand-in, a backend library named backend_name from
    backend_srcs, and one shared object per adapter linked to both."""
    inc = ["-I", STUB_DIR, "-I", os.path.join(ROOT, "include")]
    stub = os.path.join(tmp, "libmexstub.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", *inc, os.path.join(STUB_DIR, "mexstub.c"), "-o", stub],
                   check=True)
    if backend_srcs:
        subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", *inc, *backend_srcs, "-lm", "-o",
                        os.path.join(tmp, f"lib{backend_name}.so")], check=True)
    libs = {}
    for e in ENTRIES:
        out = os.path.join(tmp, f"{e}.so")
        subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra", "-Werror", *inc,
                        os.path.join(MEX_DIR, f"{e}.c"), "-L", tmp, "-lmexstub", *extra,
                        f"-Wl,-rpath,{tmp}", "-o", out], check=True)
        libs[e] = out
    return stub, libs


class Matlab:
    """Calls the adapters the way Matlab calls MEX files."""

    def __init__(self, stub_path, libs):
        self.stub = ctypes.CDLL(stub_path, mode=ctypes.RTLD_GLOBAL)
        s = self.stub
        s.stub_array.restype = ctypes.c_void_p
        s.stub_array.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int]
        s.stub_string.restype = ctypes.c_void_p
        s.stub_string.argtypes = [ctypes.c_char_p]
        s.stub_call.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        s.stub_error.restype = ctypes.c_char_p
        s.stub_warning.restype = ctypes.c_char_p
        s.stub_u64.restype = ctypes.c_uint64
        s.stub_u64.argtypes = [ctypes.c_void_p]
        self.fns = {e: ctypes.cast(ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL).mexFunction, ctypes.c_void_p)
                    for e, p in libs.items()}
        self.keep = []

    def arr(self, cls, a, complex_=0):
        a = np.ascontiguousarray(a)
        self.keep.append(a)
        return self.stub.stub_array(cls, a.size, a.ctypes.data, complex_)

    def text(self, s):
        return self.stub.stub_string(s.encode())

    def handle_arg(self, h):
        return self.arr(UINT64, np.array([h], np.uint64))

    def __call__(self, entry, *args, nlhs=0):
        prhs = (ctypes.c_void_p * max(1, len(args)))(*args)
        plhs = (ctypes.c_void_p * max(1, nlhs))()
        rc = self.stub.stub_call(self.fns[entry], nlhs, plhs, len(args), prhs)
        err = self.stub.stub_error().decode()
        out = self.stub.stub_u64(plhs[0]) if (rc == 0 and entry == "mexCreateGridder") else None
        return rc, err, out

    def create(self, n=16, p=10, f=2.0, pad=3):
        i32 = lambda v: self.arr(INT32, np.array([v], np.int32))
        return self("mexCreateGridder", i32(n), i32(p), self.arr(SINGLE, np.array([f], np.float32)), i32(pad),
                    self.arr(DOUBLE, np.array([1.0])), i32(0), self.arr(DOUBLE, np.array([1.0])),
                    self.arr(LOGICAL, np.array([0], np.uint8)), nlhs=1)


@pytest.fixture(scope="module")
def ml(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mex"))
    stub, libs = build(tmp, [os.path.join(STUB_DIR, "mock_mgg.c"),
                             os.path.join(ROOT, "multigpugridder_amd", "csrc", "mgg_mex.c")], "mockmgg",
                       extra=["-lmockmgg"])
    m = Matlab(stub, libs)
    m.mock = ctypes.CDLL(os.path.join(tmp, "libmockmgg.so"))
    m.mock.mock_log.restype = ctypes.c_char_p
    return m


def log(ml):
    return ml.mock.mock_log().decode().splitlines()


def test_create_and_delete_follow_the_handle_convention(ml):
    ml.mock.mock_reset()
    locks = ml.stub.stub_locks()
    rc, err, h = ml.create()
    assert rc == 0, err
    assert log(ml) == ["create 16 10 2 3 1 0 1 0"]
    assert ml.stub.stub_locks() == locks + 1  # mexLock while a handle exists
    rc, err, _ = ml("mexDeleteGridder", ml.handle_arg(h))
    assert rc == 0, err
    assert log(ml)[-1] == "destroy 1"
    assert ml.stub.stub_locks() == locks
    # wrong argument count, as Mex/mexCreateGridder.cpp:12-15 refuses it
    rc, err, _ = ml("mexCreateGridder", ml.arr(INT32, np.array([16], np.int32)), nlhs=1)
    assert rc == 1 and "There should be 8 inputs" in err


def test_invalid_handles_are_refused(ml):
    rc, err, _ = ml("mexMultiGPUForwardProject", ml.arr(DOUBLE, np.array([1.0])))
    assert rc == 1 and err == "Input must be a real uint64 scalar."
    junk = np.zeros(8, np.uint64)  # memory that is not a handle record
    ml.keep.append(junk)
    rc, err, _ = ml("mexMultiGPUBackProject", ml.handle_arg(junk.ctypes.data))
    assert rc == 1 and err == "Handle not valid."
    rc, err, _ = ml("mexMultiGPUForwardProject", ml.handle_arg(0))
    assert rc == 1 and err == "Handle not valid."
    # a deleted handle is refused, not read
    rc, err, h = ml.create()
    assert rc == 0, err
    assert ml("mexDeleteGridder", ml.handle_arg(h))[0] == 0
    for entry in ("mexMultiGPUForwardProject", "mexDeleteGridder"):
        rc, err, _ = ml(entry, ml.handle_arg(h))
        assert rc == 1 and err == "Handle not valid.", entry
    # an arbitrary integer is refused without being dereferenced
    rc, err, _ = ml("mexMultiGPUBackProject", ml.handle_arg(0xDEAD0000BEEF))
    assert rc == 1 and err == "Handle not valid."


def test_set_variables_reaches_each_setter(ml):
    ml.mock.mock_reset()
    rc, err, h = ml.create()
    H = lambda: ml.handle_arg(h)
    i32 = lambda *v: ml.arr(INT32, np.array(v, np.int32))
    vol = np.full((16, 16, 16), 0.5, np.float32)
    imgs1 = np.full((1, 16, 16), 1.5, np.float32)  # one image: Matlab's size() drops the trailing 1
    axes = np.arange(90, dtype=np.float32)
    table = np.linspace(0, 1, 501, dtype=np.float32)
    calls = [
        (("SetVolume", ml.arr(SINGLE, vol), i32(16, 16, 16)), "mgg_set_volume 0.5 16 16 16"),
        (("SetImages", ml.arr(SINGLE, imgs1), i32(16, 16)), "mgg_set_images 1.5 16 16 1"),
        (("SetCASImages", ml.arr(SINGLE, imgs1), i32(16, 16)), "mgg_set_cas_images 1.5 16 16 1"),
        (("SetCASVolume", ml.arr(SINGLE, vol), i32(16, 16, 16)), "mgg_set_cas_volume 0.5 16 16 16"),
        (("SetCoordAxes", ml.arr(SINGLE, axes), i32(90, 1)), "mgg_set_coord_axes 0 90 1"),
        (("SetKBTable", ml.arr(SINGLE, table), i32(1, 501)), "mgg_set_kb_table 0 501"),
        (("SetMaskRadius", ml.arr(SINGLE, np.array([7.5], np.float32))), "mgg_set_mask_radius 7.5"),
        (("SetNumAxes", i32(10)), "mgg_set_num_axes 10"),
        (("SetMaxAxesToAllocate", i32(1000)), "mgg_set_max_axes_to_allocate 1000"),
        (("SetNumberStreamsFP", i32(10)), "mgg_set_num_streams_fp 10"),
        (("SetNumberStreamsBP", i32(4)), "mgg_set_num_streams_bp 4"),
        (("SetGPUs", i32(0), i32(1)), "mgg_set_gpus 1 0"),
    ]
    for (cmd, *vals), want in calls:
        rc, err, _ = ml("mexSetVariables", ml.text(cmd), H(), *vals)
        assert rc == 0, (cmd, err)
        assert log(ml)[-1] == want, cmd
    for entry, want in (("mexMultiGPUForwardProject", "mgg_forward_project 1"),
                        ("mexMultiGPUBackProject", "mgg_back_project 1"),
                        ("mexMultiGPUGetVolume", "mgg_cas_volume_to_host 1"),
                        ("mexMultiGPUReconstructVolume", "mgg_reconstruct_volume 1")):
        rc, err, _ = ml(entry, H())
        assert rc == 0, err
        assert log(ml)[-1] == want
    ml("mexDeleteGridder", H())


def test_set_variables_rejects_bad_arguments(ml):
    rc, err, h = ml.create()
    H = lambda: ml.handle_arg(h)
    i32 = lambda *v: ml.arr(INT32, np.array(v, np.int32))
    n_before = len(log(ml))
    bad = [
        (("SetVolume", ml.arr(SINGLE, np.zeros(10, np.float32)), i32(16, 16, 16)), "holds 10 values"),
        (("SetVolume", ml.arr(DOUBLE, np.zeros(8)), i32(2, 2, 2)), "real single array"),
        (("SetVolume", ml.arr(SINGLE, np.zeros(8, np.float32), complex_=1), i32(2, 2, 2)), "real single array"),
        (("SetVolume", ml.arr(SINGLE, np.zeros(8, np.float32))), "needs the array and its size vector"),
        (("SetNumAxes", ml.arr(DOUBLE, np.array([2.5]))), "must be an integer"),
        (("SetGPUs", i32(0, 1), i32(3)), "count says 3"),
        (("Frobnicate", i32(1)), "unknown command 'Frobnicate'"),
    ]
    for (cmd, *vals), want in bad:
        rc, err, _ = ml("mexSetVariables", ml.text(cmd), H(), *vals)
        assert rc == 1 and want in err, (cmd, err)
    rc, err, _ = ml("mexSetVariables", ml.text("SetNumAxes"), H())
    assert rc == 1 and "3 or 4 inputs" in err
    rc, err, _ = ml("mexSetVariables", i32(1), H(), i32(1))
    assert rc == 1 and "command string" in err
    assert len(log(ml)) == n_before  # nothing reached the C ABI
    ml("mexDeleteGridder", H())


def test_status_codes_become_error_text_and_extra_delete_args_warn(ml):
    rc, err, h = ml.create()
    ml.mock.mock_fail_next(5)
    rc, err, _ = ml("mexMultiGPUForwardProject", ml.handle_arg(h))
    assert rc == 1 and "ForwardProject failed (status 5): mock failure" in err
    ml.mock.mock_fail_next(4)
    rc, err, _ = ml.create()
    assert rc == 1 and "mgg_create failed (status 4)" in err
    rc, err, _ = ml("mexDeleteGridder", ml.handle_arg(h), ml.arr(DOUBLE, np.array([1.0])))
    assert rc == 0 and ml.stub.stub_warning().decode() == "Delete: Unexpected arguments ignored."
