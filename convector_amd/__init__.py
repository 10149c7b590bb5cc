"""convector_amd -- Convector's per-ray path on AMD MI355X (gfx950).

Python mirror of the reference's Rust surface (ruuda/convector):

    Mesh.load / Mesh.load_with_materials     wavefront.rs:79-152
    Scene.from_meshes, Scene.camera          scene.rs:20-129
    Scene.intersect_nearest (Bvh)            scene.rs:220-244, bvh.rs:589-669
    Renderer.render_patch_u8 / accumulate_patch_f32 / new_buffer_f32 /
    buffer_f32_into_render_buffer            renderer.rs:121-410
    Batch.render                             the patch loop of main.rs:174-214,
                                             many jobs per launch

Every call goes through the C ABI of include/convector.h into the HIP
kernels of libconvector_amd.so; nothing here computes the path.
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, Optional, Sequence

import numpy as np

from . import _native
from ._native import CvError, check, lib  # noqa: F401

NO_HIT = _native.CV_NO_HIT
RNG_PIXEL = _native.CV_RNG_PIXEL
RNG_REF_PATCH = _native.CV_RNG_REF_PATCH
PTR_DEVICE = _native.CV_PTR_DEVICE
RENDER_COUNT = _native.CV_RENDER_COUNT
RENDER_TIMING = _native.CV_RENDER_TIMING
FRAMES_AUTO = 255  # cv_tuning.frames_per_step: about 2^25 paths per step
RAY_SORT_NONE = _native.CV_RAY_SORT_NONE  # cv_tuning.ray_sort_bits: queue order
RENDER_GRAPH = _native.CV_RENDER_GRAPH

NODE_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("index", "<u4"), ("len", "<u4")])
HIT_DTYPE = np.dtype([("tri_id", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
ISECT_DTYPE = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("distance", "<f4"),
                        ("material", "<u4"), ("tex_coords", "<f4", 2), ("tri_id", "<u4"),
                        ("u", "<f4"), ("v", "<f4"), ("numi_aabb", "<u4"), ("numi_tri", "<u4"),
                        ("reserved", "<u4")])


def _p(a):
    """A numpy array's data pointer, a device pointer given as int, or None."""
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return a.ctypes.data_as(ctypes.c_void_p)


def _host_buf(a, dtype, count: int, name: str):
    """Pointer of a caller buffer the library reads or writes count elements
    of: a device pointer given as int (taken as is), or a C-contiguous,
    writeable numpy array of exactly `dtype` holding at least count
    elements.  Anything else raises instead of letting the library read or
    write past the array."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if not isinstance(a, np.ndarray):
        raise TypeError("%s must be a numpy array or a device pointer (int)" % name)
    if a.dtype != np.dtype(dtype):
        raise TypeError("%s must have dtype %s, not %s" % (name, np.dtype(dtype), a.dtype))
    if not a.flags.c_contiguous or not a.flags.writeable:
        raise ValueError("%s must be C-contiguous and writeable" % name)
    if a.size < count:
        raise ValueError("%s holds %d elements, the call needs %d" % (name, a.size, count))
    return a.ctypes.data


class Mesh:
    """A triangle mesh (wavefront.rs:18-28) held by the library."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load(cls, path: str) -> "Mesh":
        return cls.load_with_materials(path, {})

    @classmethod
    def load_with_materials(cls, path: str, materials: Dict[str, int]) -> "Mesh":
        entries = (_native.MaterialEntry * max(1, len(materials)))()
        names = []
        for i, (k, v) in enumerate(materials.items()):
            names.append(k.encode())
            entries[i].name = names[-1]
            entries[i].material = int(v)
        h = ctypes.c_void_p()
        check(lib().cv_mesh_load_obj(path.encode(), ctypes.cast(entries, ctypes.c_void_p),
                                     len(materials), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_triangles(cls, tris: np.ndarray, materials: np.ndarray,
                       uv: Optional[np.ndarray] = None) -> "Mesh":
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        m = np.ascontiguousarray(materials, dtype=np.uint32).reshape(-1)
        u = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 6)
        h = ctypes.c_void_p()
        check(lib().cv_mesh_from_triangles(_p(t), _p(u), _p(m), t.shape[0], ctypes.byref(h)))
        return cls(h)

    @property
    def num_triangles(self) -> int:
        return int(lib().cv_mesh_num_triangles(self._h))

    def triangles(self):
        n = self.num_triangles
        v = np.zeros((n, 9), np.float32)
        uv = np.zeros((n, 6), np.float32)
        m = np.zeros(n, np.uint32)
        check(lib().cv_mesh_copy_triangles(self._h, _p(v), _p(uv), _p(m)))
        return v, uv, m

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().cv_mesh_destroy(h)
            self._h = None


class Camera:
    """Camera of scene.rs:20-100; changes are pushed to the scene."""

    def __init__(self, scene: "Scene"):
        self._scene = scene
        self.position = (0.0, 0.0, 0.0)
        self.position_delta = (0.0, 0.0, 0.0)
        self.orientation = (1.0, 0.0, 0.0, 0.0)
        self.orientation_delta = (0.0, 0.0, 0.0, 0.0)
        self.screen_distance = float(np.float32(1.0) / np.float32(math.sin(math.pi / 5.0)))
        self._push()

    def _push(self):
        f = lambda x: np.ascontiguousarray(x, dtype=np.float32)  # noqa: E731
        a, b, c, d = (f(self.position), f(self.position_delta), f(self.orientation),
                      f(self.orientation_delta))
        check(lib().cv_scene_set_camera(self._scene._h, _p(a), _p(b), _p(c), _p(d),
                                        float(self.screen_distance)))

    def set_position(self, position, delta=(0.0, 0.0, 0.0)):
        self.position, self.position_delta = tuple(position), tuple(delta)
        self._push()

    def set_orientation(self, orientation, delta=(0.0, 0.0, 0.0, 0.0)):
        self.orientation, self.orientation_delta = tuple(orientation), tuple(delta)
        self._push()

    def set_fov(self, fov: float):
        self.screen_distance = float(np.float32(1.0) / np.float32(math.sin(fov / 2.0)))
        self._push()

    def set_screen_distance(self, d: float):
        self.screen_distance = float(d)
        self._push()


class Scene:
    """Scene + Bvh (scene.rs:102-245, bvh.rs:32-669) resident on the GPU."""

    def __init__(self, handle):
        self._h = handle
        self.camera = Camera(self)

    @staticmethod
    def _opts(emission_override, bins=64):
        o = _native.SceneOpts()
        # 64 bins: the reference's build; 32 or 16: the same build with fewer bins
        o.builder = _native.CV_BUILDER_REFERENCE if bins == 64 else bins
        if emission_override is not None:
            o.has_emission_override = 1
            for k in range(3):
                o.emission_override[k] = float(emission_override[k])
        return o

    @classmethod
    def from_meshes(cls, meshes: Sequence[Mesh], emission_override=None, bins=64) -> "Scene":
        arr = (ctypes.c_void_p * len(meshes))(*[m._h for m in meshes])
        o = cls._opts(emission_override, bins)
        h = ctypes.c_void_p()
        check(lib().cv_scene_create(ctypes.cast(arr, ctypes.c_void_p), len(meshes),
                                    ctypes.byref(o), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_meshes_with_bvh(cls, meshes: Sequence[Mesh], nodes: np.ndarray, order: np.ndarray,
                             emission_override=None) -> "Scene":
        arr = (ctypes.c_void_p * len(meshes))(*[m._h for m in meshes])
        o = cls._opts(emission_override)
        n = np.ascontiguousarray(nodes).view(NODE_DTYPE)
        od = np.ascontiguousarray(order, dtype=np.uint32)
        h = ctypes.c_void_p()
        check(lib().cv_scene_create_with_bvh(ctypes.cast(arr, ctypes.c_void_p), len(meshes),
                                             ctypes.byref(o), _p(n), len(n), _p(od),
                                             ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_scene_data(cls, sd, with_bvh=None, bins=64) -> "Scene":
        """Scene from a convector_amd.scenes.SceneData (camera included);
        bins: centroid bins of the GPU build (64, the reference's; 32; 16)."""
        mesh = Mesh.from_triangles(sd.tris, sd.materials, sd.uv)
        if with_bvh is None:
            s = cls.from_meshes([mesh], emission_override=sd.emission_override, bins=bins)
        else:
            s = cls.from_meshes_with_bvh([mesh], with_bvh[0], with_bvh[1],
                                         emission_override=sd.emission_override)
        s.camera.set_position(sd.camera_pos, sd.camera_dpos)
        s.camera.set_orientation(sd.camera_q, sd.camera_dq)
        s.camera.set_screen_distance(sd.screen_distance)
        return s

    def set_tuning(self, stack_entries=0, extend_waves_per_cu=0, refill_threshold=0,
                   fetch_batches=0, block_waves=0, frames_per_step=0, refill_interval=0,
                   leaf_batch=0, bvh_width=0, ray_sort_bits=0, queue_parts=0):
        """Launch shape of the traversal (the autotuner's knobs)."""
        t = _native.Tuning(stack_entries, extend_waves_per_cu, refill_threshold, fetch_batches,
                           block_waves, frames_per_step, refill_interval, leaf_batch,
                           bvh_width, ray_sort_bits, queue_parts)
        check(lib().cv_scene_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = _native.Tuning()
        check(lib().cv_scene_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def bvh_info(self) -> dict:
        info = _native.BvhInfo()
        check(lib().cv_scene_get_bvh_info(self._h, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def bvh(self):
        info = self.bvh_info()
        nodes = np.zeros(info["num_nodes"], NODE_DTYPE)
        order = np.zeros(info["num_triangles"], np.uint32)
        check(lib().cv_scene_copy_bvh(self._h, _p(nodes), _p(order)))
        return nodes, order

    def intersect_nearest(self, rays: np.ndarray, full: bool = False):
        """rays: (n, 6) float32 origin+direction.  Returns the cv_hit array
        (and the cv_isect array when full)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = r.shape[0]
        hits = np.zeros(n, HIT_DTYPE)
        isects = np.zeros(n, ISECT_DTYPE) if full else None
        check(lib().cv_bvh_intersect_nearest(self._h, _p(r), n, _p(hits), _p(isects), 0))
        return (hits, isects) if full else hits

    def intersect_nearest_device(self, rays_ptr: int, n: int, hits_ptr: int, isects_ptr=None):
        check(lib().cv_bvh_intersect_nearest(self._h, ctypes.c_void_p(rays_ptr), n,
                                             ctypes.c_void_p(hits_ptr),
                                             ctypes.c_void_p(isects_ptr) if isects_ptr else None,
                                             PTR_DEVICE))

    def intersect_nearest_noalloc(self, rays_ptr: int, n: int, hits_ptr=None, isects_ptr=None):
        """cv_bvh_intersect_nearest_noalloc: device buffers only, no
        allocation (the library's fixed per-GPU scratch array)."""
        check(lib().cv_bvh_intersect_nearest_noalloc(
            self._h, ctypes.c_void_p(rays_ptr), n,
            ctypes.c_void_p(hits_ptr) if hits_ptr else None,
            ctypes.c_void_p(isects_ptr) if isects_ptr else None))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().cv_scene_destroy(h)
            self._h = None


class Renderer:
    """Renderer (renderer.rs:16-130) bound to a GPU scene."""

    def __init__(self, scene: Scene, width: int, height: int, rng_mode: int = RNG_PIXEL,
                 max_bounces: int = 5):
        self.scene = scene
        self.width, self.height = int(width), int(height)
        o = _native.RenderOpts(rng_mode, _native.CV_APPROX_PORTABLE, max_bounces)
        h = ctypes.c_void_p()
        check(lib().cv_renderer_create(scene._h, self.width, self.height, ctypes.byref(o),
                                       ctypes.byref(h)))
        self._h = h

    def new_buffer_f32(self) -> np.ndarray:
        """renderer.rs:366-374: (height/4)*(width/16) blocks of [8][3][8] f32."""
        return np.zeros(self.width * self.height * 3, np.float32)

    def new_bitmap(self) -> np.ndarray:
        return np.zeros(self.width * self.height, np.uint32)

    def _px(self, a, dtype, per_pixel, name):
        return _p(_host_buf(a, dtype, self.width * self.height * per_pixel, name))

    def render_patch_u8(self, bitmap, gbuffer, patch_width, x, y, frame_number):
        check(lib().cv_renderer_render_patch_u8(
            self._h, self._px(bitmap, np.uint32, 1, "bitmap"),
            self._px(gbuffer, np.uint32, 1, "gbuffer"), patch_width, x, y, frame_number, 0))

    def accumulate_patch_f32(self, hdr, gbuffer, patch_width, x, y, frame_number):
        check(lib().cv_renderer_accumulate_patch_f32(
            self._h, self._px(hdr, np.float32, 3, "hdr"),
            self._px(gbuffer, np.uint32, 1, "gbuffer"), patch_width, x, y, frame_number, 0))

    def render_frame_u8(self, frame_number, gbuffer=True):
        bitmap = self.new_bitmap()
        gb = self.new_bitmap() if gbuffer else None
        check(lib().cv_renderer_render_frame_u8(self._h, _p(bitmap), _p(gb), frame_number, 0))
        return bitmap, gb

    def accumulate_frames_f32(self, hdr, first_frame, n_frames, gbuffer=None, flags=0):
        check(lib().cv_renderer_accumulate_frames_f32(
            self._h, self._px(hdr, np.float32, 3, "hdr"),
            self._px(gbuffer, np.uint32, 1, "gbuffer"), first_frame, n_frames, flags))

    def new_features(self) -> np.ndarray:
        """First-hit features: 8 f32 per pixel in rows (albedo rgb, normal
        xyz, distance, count), summed over frames."""
        return np.zeros(self.width * self.height * 8, np.float32)

    def accumulate_aov_f32(self, hdr, features, first_frame, n_frames, flags=0):
        check(lib().cv_renderer_accumulate_aov_f32(
            self._h, self._px(hdr, np.float32, 3, "hdr"),
            self._px(features, np.float32, 8, "features"), first_frame, n_frames, flags))

    def accumulate_frames_device(self, hdr_ptr: int, first_frame, n_frames, flags=0):
        check(lib().cv_renderer_accumulate_frames_f32(self._h, ctypes.c_void_p(hdr_ptr), None,
                                                      first_frame, n_frames,
                                                      flags | PTR_DEVICE))

    def buffer_f32_into_render_buffer(self, hdr, num_samples) -> np.ndarray:
        bitmap = self.new_bitmap()
        check(lib().cv_renderer_buffer_f32_into_bitmap(
            self._h, self._px(hdr, np.float32, 3, "hdr"), _p(bitmap), num_samples, 0))
        return bitmap

    def stats(self) -> dict:
        s = _native.RenderStats()
        check(lib().cv_renderer_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().cv_renderer_destroy(h)
            self._h = None


class Job:
    """One render job of a batch: frames first_frame .. first_frame +
    num_frames - 1 of `scene` at width x height, accumulated into hdr (the
    16x4-block layout; a host array, or a device pointer with PTR_DEVICE)."""

    def __init__(self, scene: Scene, width: int, height: int, first_frame: int = 1,
                 num_frames: int = 1, hdr=None, gbuffer=None, features=None):
        self.scene = scene
        self.width, self.height = int(width), int(height)
        self.first_frame, self.num_frames = int(first_frame), int(num_frames)
        self.hdr = np.zeros(self.width * self.height * 3, np.float32) if hdr is None else hdr
        self.gbuffer = gbuffer
        self.features = features

    def _ptr(self, a, dtype=np.float32, per_pixel=3, name="hdr"):
        return _host_buf(a, dtype, self.width * self.height * per_pixel, name)


class Batch:
    """Many render jobs in shared launches (cv_batch)."""

    def __init__(self, max_paths_per_step: int = 0, max_bounces: int = 5):
        o = _native.RenderOpts(RNG_PIXEL, _native.CV_APPROX_PORTABLE, max_bounces)
        h = ctypes.c_void_p()
        check(lib().cv_batch_create(ctypes.byref(o), max_paths_per_step, ctypes.byref(h)))
        self._h = h

    def set_tuning(self, stack_entries=0, extend_waves_per_cu=0, refill_threshold=0,
                   fetch_batches=0, block_waves=0, frames_per_step=0, refill_interval=0,
                   leaf_batch=0, bvh_width=0, ray_sort_bits=0, queue_parts=0):
        t = _native.Tuning(stack_entries, extend_waves_per_cu, refill_threshold, fetch_batches,
                           block_waves, frames_per_step, refill_interval, leaf_batch,
                           bvh_width, ray_sort_bits, queue_parts)
        check(lib().cv_batch_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = _native.Tuning()
        check(lib().cv_batch_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    @staticmethod
    def _jobs(jobs: Sequence[Job]):
        arr = (_native.Job * max(1, len(jobs)))()
        for i, j in enumerate(jobs):
            arr[i].scene = j.scene._h
            arr[i].width, arr[i].height = j.width, j.height
            arr[i].first_frame, arr[i].num_frames = j.first_frame, j.num_frames
            arr[i].hdr = j._ptr(j.hdr)
            arr[i].gbuffer = j._ptr(j.gbuffer, np.uint32, 1, "gbuffer")
            arr[i].features = j._ptr(j.features, np.float32, 8, "features")
        return arr

    def render(self, jobs: Sequence[Job], flags: int = 0):
        check(lib().cv_batch_render(self._h, self._jobs(jobs), len(jobs), flags))

    def autotune(self, jobs: Sequence[Job], max_sample_paths: int = 0, tolerance: bool = False):
        """Times launch shapes on a sample step of these jobs and keeps the
        fastest; returns (tuning dict, step ms).  tolerance=True also sweeps
        the BVH widths 4 and 8 (hits as the reference's except at exact
        ties)."""
        t = _native.Tuning()
        ms = ctypes.c_double()
        check(lib().cv_batch_autotune(self._h, self._jobs(jobs), len(jobs), max_sample_paths,
                                      _native.CV_TUNE_TOLERANCE if tolerance else 0,
                                      ctypes.byref(t), ctypes.byref(ms)))
        return {k: getattr(t, k) for k, _ in t._fields_}, ms.value

    def stats(self) -> dict:
        s = _native.RenderStats()
        check(lib().cv_batch_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().cv_batch_destroy(h)
            self._h = None


class Denoiser:
    """The learned denoiser (cv_denoiser): a 3-level U-Net trained and run
    on the GPU's matrix cores.  Arrays are host float32 unless a device
    pointer (int) is passed with device=True."""

    NUM_INPUT = 10

    def __init__(self, seed: int = 1):
        h = ctypes.c_void_p()
        check(lib().cv_denoiser_create(seed, ctypes.byref(h)))
        self._h = h
        self.num_params = int(lib().cv_denoiser_num_params())

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        check(lib().cv_denoiser_get_params(self._h, _p(p), None))
        return p

    def grads(self) -> np.ndarray:
        g = np.zeros(self.num_params, np.float32)
        check(lib().cv_denoiser_get_params(self._h, None, _p(g)))
        return g

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        check(lib().cv_denoiser_set_params(self._h, _p(p)))

    def forward(self, inp: np.ndarray, materialize: bool = False) -> np.ndarray:
        """inp: (nb, h, w, 10) float32 -> (nb, h, w, 3) log1p radiance.
        materialize: write the upsampled decoder inputs out first, as a
        training step does (CV_DENOISE_MATERIALIZE; the same outputs)."""
        inp = np.ascontiguousarray(inp, np.float32)
        nb, h, w, c = inp.shape
        assert c == self.NUM_INPUT
        out = np.zeros((nb, h, w, 3), np.float32)
        flags = _native.CV_DENOISE_MATERIALIZE if materialize else 0
        check(lib().cv_denoiser_forward(self._h, _p(inp), _p(out), nb, h, w, flags))
        return out

    def train_step(self, inp: np.ndarray, target: np.ndarray, lr: float,
                   materialize: bool = False) -> float:
        """One Adam step on L1(log1p) against linear target (nb, h, w, 3);
        lr = 0 computes the gradients only.  Returns the loss.  materialize:
        the separate pooling / upsampling-gradient kernels
        (CV_DENOISE_MATERIALIZE; the same results)."""
        inp = np.ascontiguousarray(inp, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        nb, h, w, _ = inp.shape
        loss = ctypes.c_float()
        flags = _native.CV_DENOISE_MATERIALIZE if materialize else 0
        check(lib().cv_denoiser_train_step(self._h, _p(inp), _p(target), nb, h, w, float(lr),
                                           flags, ctypes.byref(loss)))
        return loss.value

    def train_step_device(self, inp_ptr: int, target_ptr: int, nb, h, w, lr) -> float:
        loss = ctypes.c_float()
        check(lib().cv_denoiser_train_step(self._h, ctypes.c_void_p(inp_ptr),
                                           ctypes.c_void_p(target_ptr), nb, h, w, float(lr),
                                           PTR_DEVICE, ctypes.byref(loss)))
        return loss.value

    def forward_device(self, inp_ptr: int, out_ptr: int, nb, h, w):
        check(lib().cv_denoiser_forward(self._h, ctypes.c_void_p(inp_ptr), ctypes.c_void_p(out_ptr),
                                        nb, h, w, PTR_DEVICE))

    @staticmethod
    def prepare_input(hdr, features, width, height, spp) -> np.ndarray:
        out = np.zeros((height, width, 10), np.float32)
        check(lib().cv_denoiser_prepare_input(_p(np.ascontiguousarray(hdr, np.float32)),
                                              _p(np.ascontiguousarray(features, np.float32)),
                                              width, height, spp, _p(out), 0))
        return out

    @staticmethod
    def prepare_input_device(hdr_ptr: int, features_ptr: int, width, height, spp, out_ptr: int):
        check(lib().cv_denoiser_prepare_input(ctypes.c_void_p(hdr_ptr),
                                              ctypes.c_void_p(features_ptr), width, height, spp,
                                              ctypes.c_void_p(out_ptr), PTR_DEVICE))

    @staticmethod
    def output_to_hdr(output, width, height, spp) -> np.ndarray:
        hdr = np.zeros(width * height * 3, np.float32)
        check(lib().cv_denoiser_output_to_hdr(_p(np.ascontiguousarray(output, np.float32)), width,
                                              height, spp, _p(hdr), 0))
        return hdr

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().cv_denoiser_destroy(h)
            self._h = None


class Server:
    """Request loop with continuous batching over resident scenes
    (cv_server).  Results are 16x4-block HDR buffers holding spp x radiance."""

    def __init__(self, budget_bytes: int = 0, max_paths_per_step: int = 0, max_bounces: int = 5):
        o = _native.RenderOpts(RNG_PIXEL, _native.CV_APPROX_PORTABLE, max_bounces)
        h = ctypes.c_void_p()
        check(lib().cv_server_create(ctypes.byref(o), budget_bytes, max_paths_per_step,
                                     ctypes.byref(h)))
        self._h = h
        self._scenes = []
        self._denoiser = None
        self._live = {}

    def add_scene(self, scene: Scene) -> int:
        i = ctypes.c_uint32()
        check(lib().cv_server_add_scene(self._h, scene._h, ctypes.byref(i)))
        self._scenes.append(scene)
        return i.value

    def set_denoiser(self, d: "Denoiser"):
        check(lib().cv_server_set_denoiser(self._h, d._h if d else None))
        self._denoiser = d

    def set_tuning(self, stack_entries=0, extend_waves_per_cu=0, refill_threshold=0,
                   fetch_batches=0, block_waves=0, frames_per_step=0, refill_interval=0,
                   leaf_batch=0, bvh_width=0, ray_sort_bits=0, queue_parts=0):
        t = _native.Tuning(stack_entries, extend_waves_per_cu, refill_threshold, fetch_batches,
                           block_waves, frames_per_step, refill_interval, leaf_batch,
                           bvh_width, ray_sort_bits, queue_parts)
        check(lib().cv_server_set_tuning(self._h, ctypes.byref(t)))

    def submit(self, scene_id: int, width: int, height: int, spp: int, camera_pos, camera_q,
               denoise: bool = False, screen_distance: float = 0.0, hdr=None):
        """Queues a request; returns (ticket, hdr) -- hdr is filled when done."""
        if hdr is None:
            hdr = np.zeros(width * height * 3, np.float32)
        r = _native.Request()
        r.scene, r.width, r.height, r.spp = scene_id, width, height, spp
        r.denoise = 1 if denoise else 0
        r.flags = PTR_DEVICE if isinstance(hdr, int) else 0
        r.camera_pos[:] = [float(x) for x in camera_pos]
        r.camera_q[:] = [float(x) for x in camera_q]
        r.screen_distance = float(screen_distance)
        r.hdr = _host_buf(hdr, np.float32, width * height * 3, "hdr")
        t = ctypes.c_uint64()
        check(lib().cv_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        self._live[t.value] = hdr
        return t.value, hdr

    def poll(self, ticket: int) -> bool:
        d = ctypes.c_int()
        check(lib().cv_server_poll(self._h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int) -> float:
        """Blocks until the request is delivered; returns its latency (ms)."""
        ms = ctypes.c_double()
        try:
            check(lib().cv_server_wait(self._h, ticket, ctypes.byref(ms)))
        finally:
            self._live.pop(ticket, None)
        return ms.value

    def stats(self) -> dict:
        s = _native.ServerStats()
        check(lib().cv_server_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def reset_stats(self):
        check(lib().cv_server_reset_stats(self._h))

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().cv_server_destroy(h)
            self._h = None
            self._live.clear()

    def __del__(self):
        self.close()


def trace_enable(limit: int = 4096):
    """Start recording device kernel spans (trace.rs TraceLog)."""
    check(lib().cv_trace_enable(limit))


def trace_disable():
    check(lib().cv_trace_disable())


def trace_export(path: str):
    """Write the recorded spans as a chrome://tracing JSON file."""
    check(lib().cv_trace_export(path.encode()))


def device_synchronize():
    check(lib().cv_device_synchronize())


def version() -> str:
    return lib().cv_version().decode()
