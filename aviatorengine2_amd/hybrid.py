"""Host side of the hybrid renderer's frame sequence.

Renderer::RenderFrame in hybrid mode (renderer/renderer.cpp:1330-1583) keeps
per-renderer state - frame index, accumulation frame, pending history reset
or camera move, last frame's view-projection - and writes it into the frame's
Config before the G-buffer and compute passes run. HybridFrames does the same
for avrt_render_hybrid:

    frames = HybridFrames(renderer, config, trace_scale=0.5)
    frames.set_camera(source, target)      # CameraMoved (renderer.h:357-367)
    cfg, stats = frames.render()           # one RenderFrame

Camera matrices follow math/mat4.h:56-94 (LookAtRH, PerspectiveRH_ZO; row
major, float32 arithmetic in the reference's operation order), the jitter
renderer.cpp:1379-1391 with Renderer::Halton (avrt_halton).
"""
from __future__ import annotations

import math

import numpy as np

from . import abi

f32 = np.float32


def _normalize(v):
    # Vec3 normalize: (1 / length) * v, length = sqrt(x*x + y*y + z*z)
    ln = f32(math.sqrt(f32(f32(v[0] * v[0]) + f32(v[1] * v[1])) + f32(v[2] * v[2])))
    inv = f32(f32(1.0) / ln)
    return [f32(inv * v[0]), f32(inv * v[1]), f32(inv * v[2])]


def _cross(a, b):
    return [f32(a[1] * b[2] - a[2] * b[1]), f32(a[2] * b[0] - a[0] * b[2]), f32(a[0] * b[1] - a[1] * b[0])]


def _dot(a, b):
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def look_at_rh(eye, target, up) -> np.ndarray:
    """LookAtRH (math/mat4.h:56-69)."""
    eye = [f32(x) for x in eye]
    target = [f32(x) for x in target]
    up = [f32(x) for x in up]
    w = _normalize([f32(eye[i] - target[i]) for i in range(3)])
    u = _normalize(_cross(up, w))
    v = _cross(w, u)
    r = np.eye(4, dtype=np.float32)
    r[0, :3], r[0, 3] = u, -_dot(u, eye)
    r[1, :3], r[1, 3] = v, -_dot(v, eye)
    r[2, :3], r[2, 3] = w, -_dot(w, eye)
    return r


def perspective_rh_zo(fov_deg, aspect, z_near, z_far) -> np.ndarray:
    """PerspectiveRH_ZO (math/mat4.h:83-94). The tangent is evaluated in
    double and rounded to float."""
    f = f32(f32(1.0) / f32(math.tan(float(f32(f32(fov_deg) * f32(0.017453292519943295)) * f32(0.5)))))
    r = np.zeros((4, 4), np.float32)
    r[0, 0] = f32(f / f32(aspect))
    r[1, 1] = f
    r[2, 2] = f32(f32(z_far) / f32(f32(z_near) - f32(z_far)))
    r[2, 3] = f32(f32(f32(z_near) * f32(z_far)) / f32(f32(z_near) - f32(z_far)))
    r[3, 2] = f32(-1.0)
    return r


def mat_mul(a, b) -> np.ndarray:
    """operator*(Mat4, Mat4) (math/mat4.h:24-31) with float32 sums left to right."""
    r = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            s = f32(a[i, 0] * b[0, j])
            s = f32(s + f32(a[i, 1] * b[1, j]))
            s = f32(s + f32(a[i, 2] * b[2, j]))
            s = f32(s + f32(a[i, 3] * b[3, j]))
            r[i, j] = s
    return r


class HybridFrames:
    """Frame bookkeeping of Renderer in hybrid mode. `renderer` is a
    renderer.Renderer with the scene loaded; `config` an AvrtConfig record
    with the scene's camera, sky, size, samples and depth."""

    def __init__(self, renderer, config, trace_scale=0.5, history_while_moving=32, z_near=0.1, z_far=10000.0,
                 max_samples=0, debug_view=0):
        self.r = renderer
        self.cfg = np.array(config, dtype=abi.CONFIG_DTYPE, copy=True).reshape(1)
        self.cfg["zNear"] = z_near
        self.cfg["zFar"] = z_far
        self.cfg["debugView"] = debug_view
        self.trace_scale = float(trace_scale)
        self.history_while_moving = max(int(history_while_moving), 1)
        self.max_samples = int(max_samples)
        self.frame_index = 0
        self.accumulation_frame = 0
        self.history_reset_pending = True
        self.camera_moved_pending = False
        self.previous_view_proj = np.eye(4, dtype=np.float32)
        self.view_proj = self._camera_matrices()

    # -- Renderer state changes (renderer.h:345-367) --------------------------
    def reset_accumulation(self):
        self.accumulation_frame = 0
        self.history_reset_pending = True

    def camera_moved(self):
        self.accumulation_frame = 0
        self.camera_moved_pending = True

    def set_camera(self, source, target):
        c = self.cfg
        c["source_x"], c["source_y"], c["source_z"] = source
        c["target_x"], c["target_y"], c["target_z"] = target
        self.view_proj = self._camera_matrices()
        self.camera_moved()

    def _camera_matrices(self):
        c = self.cfg
        eye = (c["source_x"][0], c["source_y"][0], c["source_z"][0])
        tgt = (c["target_x"][0], c["target_y"][0], c["target_z"][0])
        up = (c["up_x"][0], c["up_y"][0], c["up_z"][0])
        aspect = f32(f32(c["width"][0]) / f32(c["height"][0]))
        view = look_at_rh(eye, tgt, up)
        proj = perspective_rh_zo(c["fov"][0], aspect, c["zNear"][0], c["zFar"][0])
        return mat_mul(proj, view)

    def converged(self) -> bool:
        return (self.max_samples > 0 and int(self.cfg["debugView"][0]) == 0 and
                self.accumulation_frame * int(self.cfg["samples"][0]) >= self.max_samples)

    # -- RenderFrame (renderer.cpp:1333-1423) ---------------------------------
    def next_config(self):
        c = self.cfg.copy()
        conv = self.converged()
        final = int(c["debugView"][0]) == 0
        c["batch"] = self.accumulation_frame
        c["displayOnly"] = 1 if conv else 0
        c["frameIndex"] = self.frame_index
        c["historyReset"] = 1 if self.history_reset_pending else 0
        c["cameraMoved"] = 1 if self.camera_moved_pending else 0
        c["historyCap"] = float(self.history_while_moving)
        c["prevViewProj"][0] = self.previous_view_proj.reshape(16)
        c["jitter_x"] = 0.0
        c["jitter_y"] = 0.0
        if not conv and final:
            index = (self.frame_index % 1024) + 1
            c["jitter_x"] = f32(f32(self.r.lib.avrt_halton(index, 2)) - f32(0.5))
            c["jitter_y"] = f32(f32(self.r.lib.avrt_halton(index, 3)) - f32(0.5))
        c["traceScale"] = self.trace_scale if final else 1.0
        return c

    def advance(self, config):
        """State update after a frame (renderer.cpp:1573-1582)."""
        if int(config["displayOnly"][0]) == 0:
            self.accumulation_frame += 1
            self.frame_index += 1
            self.history_reset_pending = False
            self.camera_moved_pending = False
            self.previous_view_proj = self.view_proj.copy()

    def render(self):
        cfg = self.next_config()
        stats = self.r.RenderHybrid(cfg)
        self.advance(cfg)
        return cfg, stats


def orbit(center, radius, height, degrees):
    """Camera position on a circle around `center` (y up)."""
    a = math.radians(degrees)
    return (center[0] + radius * math.sin(a), center[1] + height, center[2] + radius * math.cos(a))
