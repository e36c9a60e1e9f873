"""The C ABI: byte layouts match the reference's static_asserts, and
libavrt.so exports every entry point include/avrt.h declares. No GPU calls."""
import ctypes
import os
import re

import numpy as np

from aviatorengine2_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_object_layout_matches_reference_static_asserts():
    # scene/gpu_types.h:104-112
    f = abi.OBJECT_DTYPE.fields
    assert abi.OBJECT_DTYPE.itemsize == 128
    assert f["radius"][1] == 28
    assert f["rotation"][1] == 76
    assert f["density"][1] == 80
    assert f["emission"][1] == 84
    assert f["textureTint"][1] == 92
    assert f["r"][1] == 96
    assert f["fuzz"][1] == 108
    assert f["textureID"][1] == 124
    assert f["instanceIndex"][1] == 88
    assert f["shapeType"][1] == 116 and f["colorType"][1] == 120


def test_config_layout_matches_reference_static_asserts():
    # renderer/renderer.h:151-159
    f = abi.CONFIG_DTYPE.fields
    assert abi.CONFIG_DTYPE.itemsize == 240
    assert f["exposure"][1] == 124
    assert f["debugView"][1] == 128
    assert f["jitter_y"][1] == 144
    assert f["displayOnly"][1] == 148
    assert f["traceScale"][1] == 152
    assert f["frameIndex"][1] == 156
    assert f["historyCap"][1] == 168
    assert f["prevViewProj"][1] == 176
    # cbuffer offsets the compute shaders read (path_trace.comp.hlsl:118-148)
    assert f["up_x"][1] == 48 and f["sky_r"][1] == 64 and f["width"][1] == 76
    assert f["horizon_r"][1] == 80 and f["height"][1] == 92 and f["samples"][1] == 96
    assert f["batch"][1] == 104 and f["depth"][1] == 108 and f["num_lights"][1] == 120


def test_node_instance_vertex_layouts():
    assert abi.NODE_DTYPE.itemsize == 48  # scene/bvh.h:43
    assert abi.NODE_DTYPE.fields["left"][1] == 28 and abi.NODE_DTYPE.fields["count"][1] == 36
    assert abi.INSTANCE_DTYPE.itemsize == 48  # gpu_types.h:96
    assert abi.INSTANCE_DTYPE.fields["qx"][1] == 16 and abi.INSTANCE_DTYPE.fields["vertexBase"][1] == 32
    assert abi.INSTANCE_DTYPE.fields["blasBase"][1] == 44
    assert abi.VERTEX_DTYPE.itemsize == 32  # mesh_library.h:46


def test_oracle_layouts_agree(orc):
    assert orc.orc_sizeof_config() == 240


def _declared_in_header():
    text = open(os.path.join(ROOT, "include", "avrt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(avrt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(avrt_lib):
    declared = _declared_in_header()
    assert declared, "no declarations found"
    missing = [s for s in declared if not hasattr(avrt_lib, s)]
    assert not missing, missing
    assert sorted(abi.declared_symbols()) == declared


def test_create_without_gpu_fails_cleanly_or_succeeds(avrt_lib):
    """avrt_create never throws across the ABI: on a host without a GPU it
    returns AVRT_E_DEVICE with a message."""
    import torch  # noqa: F401  (only to know whether a GPU is visible)
    if torch.cuda.is_available():
        return
    ctx = ctypes.c_void_p()
    rc = avrt_lib.avrt_create(None, ctypes.byref(ctx))
    assert rc == -4
    assert b"device" in avrt_lib.avrt_last_error(None)


def test_create_rejects_unknown_trace_schedule(avrt_lib):
    """Options are checked before any device call."""
    o = abi.AvrtCreateOptions()
    o.trace_schedule = 3
    ctx = ctypes.c_void_p()
    assert avrt_lib.avrt_create(ctypes.byref(o), ctypes.byref(ctx)) == -1
    assert b"trace_schedule" in avrt_lib.avrt_last_error(None)
    assert not ctx.value


def test_version_string(avrt_lib):
    assert avrt_lib.avrt_version().startswith(b"avrt")


def test_host_builders_reject_bad_arguments(avrt_lib):
    cnt = ctypes.c_uint32(7)
    assert avrt_lib.avrt_build_bvh(None, 0, None, ctypes.byref(cnt), None) == 0 and cnt.value == 0
    assert avrt_lib.avrt_build_bvh(None, 3, None, ctypes.byref(cnt), None) == -1
    pos = np.zeros((3, 3), np.float32)
    tri = np.array([[0, 1, 5]], np.uint32)  # index out of range
    nodes = np.zeros(2, abi.NODE_DTYPE)
    idx = np.zeros(3, np.uint32)
    local = np.zeros(1, abi.AABB_DTYPE)
    rc = avrt_lib.avrt_build_blas(abi.ptr(pos), 3, abi.ptr(tri), 1, abi.ptr(nodes), ctypes.byref(cnt), abi.ptr(idx),
                                  abi.ptr(local))
    assert rc == -1
