"""The device-pointer operators reject tensors whose sizes do not match what
their kernels read and write, before any pointer reaches a kernel (CPU: the
shape checks come before the device checks, so CPU tensors exercise them)."""
import pytest
import torch

from multigpugridder_amd import ops

f32 = torch.float32


def z(*shape):
    return torch.zeros(shape, dtype=f32)


def rejects(fn, *args, match, **kw):
    with pytest.raises(ValueError, match=match):
        fn(*args, **kw)


def test_forward_project_shapes():
    cas, axes = z(38, 38, 38), z(4, 9)
    rejects(ops.forward_project, z(38, 38, 37), axes, 32, match="cube")
    rejects(ops.forward_project, z(38 ** 3), axes, 32, match="cube")
    rejects(ops.forward_project, cas, z(4, 8), 32, match="9 floats")
    rejects(ops.forward_project, cas, axes, 32, out=z(3, 32, 32), match="out must hold exactly 4096")
    rejects(ops.forward_project, cas, axes, 0, match="positive")
    # well-formed shapes get as far as the device check
    rejects(ops.forward_project, cas, axes, 32, out=z(4, 32, 32), match="device tensor")


def test_back_project_shapes():
    imgs, axes = z(4, 32, 32), z(4, 9)
    rejects(ops.back_project, z(4, 32, 31), axes, 38, match="square")
    rejects(ops.back_project, z(3, 32, 32), axes, 38, match="3 images for 4")
    rejects(ops.back_project, imgs, z(37), 38, match="9 floats")
    rejects(ops.back_project, imgs, axes, 38, out=z(37, 38, 38), match="out must hold exactly")
    rejects(ops.back_project, imgs, axes, 38, out=z(38, 38, 38), density=z(10), match="density must hold")
    rejects(ops.back_project, imgs, axes, 38, out=z(38, 38, 38), density=z(38, 38, 38), match="device tensor")
    rejects(ops.back_project_adjoint, imgs, axes, 38, out=z(38, 38, 39), match="out must hold exactly")
    rejects(ops.back_project_adjoint, z(2, 32, 32), axes, 38, match="2 images for 4")


def test_chain_shapes():
    rejects(ops.cas_images_to_images, z(2, 32, 30), 16, match="square")
    rejects(ops.cas_images_to_images, z(2, 32, 32), 33, match="n must be in")
    rejects(ops.cas_images_to_images, z(2, 32, 32), 16, out=z(2, 16, 15), match="out must hold exactly")
    rejects(ops.images_to_cas_images, z(2, 16, 16), 8, match="m must be >= 16")
    rejects(ops.images_to_cas_images, z(2, 16, 16), 32, out=z(2, 32, 31), match="out must hold exactly")
    rejects(ops.cas_images_to_images_transpose, z(2, 16, 16), 32, out=z(1, 32, 32), match="out must hold")
    rejects(ops.volume_to_cas_volume, z(16, 16, 15), match="cube")
    rejects(ops.volume_to_cas_volume, z(16, 16, 16), out=z(38, 38, 37), match="out must hold exactly")
    rejects(ops.cas_volume_to_volume, z(37, 38, 38), 16, match="cas_volume must hold exactly")
    rejects(ops.cas_volume_to_volume, z(38, 38, 38), 16, out=z(16, 16, 15), match="out must hold exactly")
    rejects(ops.volume_to_cas_volume_transpose, z(38, 38, 37), 16, match="cas_volume must hold")


def test_elementwise_shapes():
    rejects(ops.divide_density, z(8), z(7), match="den must hold exactly 8")
    rejects(ops.divide_density, z(8), z(8), out=z(9), match="out must hold exactly 8")
    rejects(ops.reduce_partials, z(0, 8), match="K >= 1")
    rejects(ops.reduce_partials, z(3, 8), out=z(7), match="out must hold exactly 8")
    rejects(ops.divide_density, z(8), z(8), match="device tensor")
