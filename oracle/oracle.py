"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check of
__graft_entry__ and the cpu_baseline leg of bench.py, and only as the checker
or the timed CPU baseline.  The product (convector_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

APPROX_PORTABLE = 0
APPROX_HOST_X86 = 1
RNG_PIXEL = 0
RNG_REF_PATCH = 1
NO_HIT = 0xFFFFFFFF

NODE_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("index", "<u4"), ("len", "<u4")])

_lib = None


def build(force: bool = False) -> str:
    """Compiles the oracle with its Makefile (gcc only)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
U32 = ctypes.c_uint32
F32 = ctypes.c_float
I32 = ctypes.c_int
SZ = ctypes.c_size_t


def _declare(L):
    sig = {
        "or_rcp": (F32, [F32, I32]),
        "or_rsqrt": (F32, [F32, I32]),
        "or_sin_poly": (F32, [F32]),
        "or_cos_poly": (F32, [F32]),
        "or_acos_poly": (F32, [F32]),
        "or_acos_fast_poly": (F32, [F32]),
        "or_recip_precise": (F32, [F32, I32]),
        "or_cvt_f32_i32": (ctypes.c_int32, [F32]),
        "or_x86_min": (F32, [F32, F32]),
        "or_x86_max": (F32, [F32, F32]),
        "or_geq_mask": (U32, [F32, F32]),
        "or_rng_ref_draws": (None, [U32, U32, U32, U32, P]),
        "or_rng_ref_floats": (None, [U32, U32, U32, U32, I32, P]),
        "or_rng_ref_hemisphere": (None, [U32, U32, U32, U32, P]),
        "or_rng_pixel_draws": (None, [U32, U32, U32, P]),
        "or_rotate_hemisphere": (None, [P, P, SZ, I32, P]),
        "or_quat_rotate": (None, [P, P, SZ, P]),
        "or_quat_interpolate": (None, [P, P, P, SZ, P]),
        "or_aabb_intersect": (None, [P, P, SZ, I32, P, P, P]),
        "or_aabb_area": (F32, [P]),
        "or_triangle_intersect": (I32, [P, P, F32, P]),
        "or_triangle_intersect_direct": (None, [P, P, I32, P]),
        "or_bvh_build": (I32, [P, U32, P, P, P, P, P]),
        "or_bvh_build_bins": (I32, [P, U32, U32, P, P, P, P, P]),
        "or_tree_sah_cost": (F32, [P, P, U32]),
        "or_scene_create": (P, [P, P, P, U32, P, U32, P, I32, P]),
        "or_scene_destroy": (None, [P]),
        "or_scene_set_camera": (None, [P, P, P, P, P, F32]),
        "or_scene_set_emission_override": (None, [P, I32, P]),
        "or_scene_set_approx": (None, [P, I32]),
        "or_scene_num_nodes": (U32, [P]),
        "or_scene_num_direct": (U32, [P]),
        "or_scene_copy_bvh": (None, [P, P, P]),
        "or_intersect_nearest": (None, [P, P, SZ, P, P, P, P, P, I32]),
        "or_intersect_nearest_packet": (None, [P, P, SZ, P, P, P, P]),
        "or_intersect_brute_f64": (None, [P, P, SZ, P, P]),
        "or_render_patch_u8": (I32, [P, U32, U32, P, P, U32, U32, U32, U32, I32]),
        "or_accumulate_patch_f32": (I32, [P, U32, U32, P, P, U32, U32, U32, U32, I32]),
        "or_buffer_f32_into_bitmap": (I32, [U32, U32, P, P, U32]),
        "or_accumulate_frames": (I32, [P, U32, U32, P, P, U32, U32, I32, I32]),
        "or_render_frame_u8": (I32, [P, U32, U32, P, P, U32, I32, I32]),
        "or_frame_work": (I32, [P, U32, U32, U32, I32, I32, P, P, P]),
        "or_hdr_blocks_to_rows": (None, [U32, U32, P, P]),
        "or_sample_patches": (I32, [P, U32, U32, U32, I32, I32, P, U32, P]),
        "or_avx_available": (I32, []),
        "or_avx_accumulate_frames": (I32, [P, U32, U32, P, U32, U32, I32, I32, P]),
        "or_avx_sample_patches": (I32, [P, U32, U32, U32, I32, P, U32, P]),
        "or_avx_intersect_packets": (I32, [P, P, SZ, P, P, I32]),
        "or_microfacet_normal_dist": (F32, [P, P, ctypes.c_int32]),
        "or_microfacet_fresnel": (None, [P, P, P, P, P]),
        "or_microfacet_brdf": (None, [U32, P, P, P, I32, I32, P, P]),
        "or_pd_brdf": (F32, [P, P]),
        "or_sky_intensity_lane": (None, [P, P]),
        "or_pd_direct": (I32, [P, P, SZ, P]),
        "or_direct_sample": (I32, [P, P, SZ, P, P, P]),
        "or_continue_path": (I32, [P, SZ, P, P, P, I32, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def bvh_build(tris: np.ndarray, bins: int = 64):
    """Reference BVH (bvh.rs:487-545); bins 16 or 32: the same build with
    that many centroid bins.  Returns (status, nodes, order, stats)."""
    tris = f32(tris)
    n = tris.shape[0]
    nodes = np.zeros(max(2 * n, 2), dtype=NODE_DTYPE)
    order = np.zeros(max(n, 1), dtype=np.uint32)
    nn = np.zeros(1, dtype=np.uint32)
    ar = np.zeros(1, dtype=np.float32)
    tl = np.zeros(1, dtype=np.float32)
    st = lib().or_bvh_build_bins(ptr(tris), n, bins, ptr(nodes), ptr(nn), ptr(order), ptr(ar),
                                 ptr(tl))
    return st, nodes[: int(nn[0])].copy(), order[:n].copy(), (float(ar[0]), float(tl[0]))


class Scene:
    """Oracle scene: BVH + triangles in BVH order + camera."""

    def __init__(self, tris, materials, uv=None, nodes=None, order=None,
                 approx=APPROX_PORTABLE):
        self.tris = f32(tris)
        self.materials = u32(materials)
        self.uv = None if uv is None else f32(uv)
        st = np.zeros(1, dtype=np.int32)
        nodes_c = None if nodes is None else np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        order_c = None if order is None else u32(order)
        nn = 0 if nodes is None else len(nodes_c)
        self._h = lib().or_scene_create(ptr(self.tris), ptr(self.uv), ptr(self.materials),
                                        self.tris.shape[0], ptr(nodes_c), nn, ptr(order_c),
                                        approx, ptr(st))
        if not self._h:
            raise RuntimeError("oracle scene creation failed with status %d" % int(st[0]))
        self.ntris = self.tris.shape[0]

    @classmethod
    def from_scene_data(cls, sd, approx=APPROX_PORTABLE, nodes=None, order=None, bins=64):
        if nodes is None and bins != 64:
            st, nodes, order, _ = bvh_build(sd.tris, bins)
            if st != 0:
                raise RuntimeError("oracle %d-bin build failed with status %d" % (bins, st))
        s = cls(sd.tris, sd.materials, sd.uv, nodes=nodes, order=order, approx=approx)
        s.set_camera(*sd.camera_args())
        if sd.emission_override is not None:
            s.set_emission_override(sd.emission_override)
        return s

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_scene_destroy(h)
            self._h = None

    def set_camera(self, pos, dpos, q, dq, screen_distance):
        lib().or_scene_set_camera(self._h, ptr(f32(pos)), ptr(f32(dpos)), ptr(f32(q)),
                                  ptr(f32(dq)), float(screen_distance))

    def set_emission_override(self, rgb):
        if rgb is None:
            lib().or_scene_set_emission_override(self._h, 0, None)
        else:
            lib().or_scene_set_emission_override(self._h, 1, ptr(f32(rgb)))

    def set_approx(self, approx):
        lib().or_scene_set_approx(self._h, approx)

    @property
    def num_direct(self):
        return lib().or_scene_num_direct(self._h)

    def bvh(self):
        nn = lib().or_scene_num_nodes(self._h)
        nodes = np.zeros(nn, dtype=NODE_DTYPE)
        order = np.zeros(self.ntris, dtype=np.uint32)
        lib().or_scene_copy_bvh(self._h, ptr(nodes), ptr(order))
        return nodes, order

    def intersect(self, rays, nthreads=8, isect=False):
        rays = f32(rays)
        n = rays.shape[0]
        tri = np.zeros(n, dtype=np.uint32)
        tuv = np.zeros((n, 3), dtype=np.float32)
        na = np.zeros(n, dtype=np.uint32)
        nt = np.zeros(n, dtype=np.uint32)
        full = np.zeros((n, 12), dtype=np.float32) if isect else None
        lib().or_intersect_nearest(self._h, ptr(rays), n, ptr(tri), ptr(tuv), ptr(na), ptr(nt),
                                   ptr(full), nthreads)
        out = {"tri": tri, "tuv": tuv, "numi_aabb": na, "numi_tri": nt}
        if isect:
            out["isect"] = full
        return out

    def intersect_packet(self, rays):
        rays = f32(rays)
        n = rays.shape[0]
        assert n % 8 == 0
        tri = np.zeros(n, dtype=np.uint32)
        tuv = np.zeros((n, 3), dtype=np.float32)
        na = np.zeros(n, dtype=np.uint32)
        nt = np.zeros(n, dtype=np.uint32)
        lib().or_intersect_nearest_packet(self._h, ptr(rays), n, ptr(tri), ptr(tuv), ptr(na),
                                          ptr(nt))
        return {"tri": tri, "tuv": tuv, "numi_aabb": na, "numi_tri": nt}

    def intersect_brute_f64(self, rays):
        rays = f32(rays)
        n = rays.shape[0]
        tri = np.zeros(n, dtype=np.uint32)
        tuv = np.zeros((n, 3), dtype=np.float64)
        lib().or_intersect_brute_f64(self._h, ptr(rays), n, ptr(tri), ptr(tuv))
        return {"tri": tri, "tuv": tuv}

    def accumulate_frames(self, width, height, first_frame, n_frames, rng_mode=RNG_PIXEL,
                          nthreads=8, hdr=None, gbuffer=None):
        if hdr is None:
            hdr = np.zeros(width * height * 3, dtype=np.float32)
        if gbuffer is None:
            gbuffer = np.zeros(width * height, dtype=np.uint32)
        st = lib().or_accumulate_frames(self._h, width, height, ptr(hdr), ptr(gbuffer),
                                        first_frame, n_frames, rng_mode, nthreads)
        if st:
            raise RuntimeError("oracle accumulate_frames status %d" % st)
        return hdr, gbuffer

    def render_frame_u8(self, width, height, frame, rng_mode=RNG_PIXEL, nthreads=8):
        bitmap = np.zeros(width * height, dtype=np.uint32)
        gbuffer = np.zeros(width * height, dtype=np.uint32)
        st = lib().or_render_frame_u8(self._h, width, height, ptr(bitmap), ptr(gbuffer), frame,
                                      rng_mode, nthreads)
        if st:
            raise RuntimeError("oracle render_frame_u8 status %d" % st)
        return bitmap, gbuffer

    def render_patch_u8(self, width, height, bitmap, gbuffer, patch_width, x, y, frame,
                        rng_mode=RNG_REF_PATCH):
        return lib().or_render_patch_u8(self._h, width, height, ptr(bitmap), ptr(gbuffer),
                                        patch_width, x, y, frame, rng_mode)

    def accumulate_patch_f32(self, width, height, hdr, gbuffer, patch_width, x, y, frame,
                             rng_mode=RNG_REF_PATCH):
        return lib().or_accumulate_patch_f32(self._h, width, height, ptr(hdr), ptr(gbuffer),
                                             patch_width, x, y, frame, rng_mode)

    def frame_work(self, width, height, frame, rng_mode=RNG_PIXEL, nthreads=8):
        q = np.zeros(1, dtype=np.uint64)
        a = np.zeros(1, dtype=np.uint64)
        t = np.zeros(1, dtype=np.uint64)
        st = lib().or_frame_work(self._h, width, height, frame, rng_mode, nthreads, ptr(q),
                                 ptr(a), ptr(t))
        if st:
            raise RuntimeError("oracle frame_work status %d" % st)
        return int(q[0]), int(a[0]), int(t[0])

    def sample_patches(self, width, height, frame, patch_ids, rng_mode=RNG_REF_PATCH,
                       nthreads=8):
        ids = u32(patch_ids)
        out = np.zeros(3, dtype=np.uint64)
        st = lib().or_sample_patches(self._h, width, height, frame, rng_mode, nthreads, ptr(ids),
                                     len(ids), ptr(out))
        if st:
            raise RuntimeError("oracle sample_patches status %d" % st)
        return int(out[0]), int(out[1]), int(out[2])

    def pd_direct(self, rays):
        """pd_direct_sample (material.rs:281-320) for n rays (n x 6)."""
        rays = f32(rays)
        out = np.zeros(rays.shape[0], dtype=np.float32)
        if lib().or_pd_direct(self._h, ptr(rays), rays.shape[0], ptr(out)):
            raise RuntimeError("scene has no direct-sample triangles")
        return out

    def direct_sample(self, draws3):
        """Scene::get_direct_sample (scene.rs:143-202) per lane from its
        (sample_u32, sample_unit, sample_unit) draws: (position, normal, area)."""
        d = u32(draws3).reshape(-1, 3)
        n = d.shape[0]
        pos = np.zeros((n, 3), dtype=np.float32)
        nrm = np.zeros((n, 3), dtype=np.float32)
        area = np.zeros(n, dtype=np.float32)
        if lib().or_direct_sample(self._h, ptr(d), n, ptr(pos), ptr(nrm), ptr(area)):
            raise RuntimeError("direct sampling needs exactly 8 direct-sample triangles")
        return pos, nrm, area

    def continue_path(self, ray7, isect7, draws6, ignore_fresnel=False):
        """continue_path (material.rs:327-411) for independent lanes:
        returns (new ray n x 7, colour modulation n x 3, raw Fresnel n)."""
        ray7 = f32(ray7).reshape(-1, 7)
        isect7 = f32(isect7).reshape(-1, 7)
        d = u32(draws6).reshape(-1, 6)
        n = ray7.shape[0]
        assert isect7.shape[0] == n and d.shape[0] == n
        out = np.zeros((n, 7), dtype=np.float32)
        cm = np.zeros((n, 3), dtype=np.float32)
        fr = np.zeros(n, dtype=np.float32)
        if lib().or_continue_path(self._h, n, ptr(ray7), ptr(isect7), ptr(d),
                                  1 if ignore_fresnel else 0, ptr(out), ptr(cm), ptr(fr)):
            raise RuntimeError("continue_path needs exactly 8 direct-sample triangles")
        return out, cm, fr

    def avx_accumulate_frames(self, width, height, first_frame, n_frames, per_lane=False,
                              nthreads=8, hdr=None):
        """The 8-wide AVX2 path (reference Rng per patch, the scene's approx
        mode) with the reference's packet traversal, or with each lane
        traversed alone (per_lane): returns (hdr in the 16x4 block layout,
        extend queries)."""
        if hdr is None:
            hdr = np.zeros(width * height * 3, dtype=np.float32)
        q = np.zeros(1, dtype=np.uint64)
        st = lib().or_avx_accumulate_frames(self._h, width, height, ptr(hdr), first_frame,
                                            n_frames, 1 if per_lane else 0, nthreads, ptr(q))
        if st:
            raise RuntimeError("oracle avx_accumulate_frames status %d" % st)
        return hdr, int(q[0])

    def avx_sample_patches(self, width, height, frame, patch_ids, nthreads=8):
        """The AVX2 packet path over the listed 32x32 patches of one frame;
        returns the number of extend queries."""
        ids = u32(patch_ids)
        q = np.zeros(1, dtype=np.uint64)
        st = lib().or_avx_sample_patches(self._h, width, height, frame, nthreads, ptr(ids),
                                         len(ids), ptr(q))
        if st:
            raise RuntimeError("oracle avx_sample_patches status %d" % st)
        return int(q[0])

    def avx_intersect_packets(self, rays, nthreads=8):
        """Bvh::intersect_nearest on packets of 8 consecutive rays with AVX2
        (bvh.rs:589-657; the bench shape of bvh.rs:681-739).  Returns
        (source triangle index or 0xffffffff, distance) per ray."""
        rays = f32(rays)
        n = rays.shape[0]
        tri = np.zeros(n, dtype=np.uint32)
        dist = np.zeros(n, dtype=np.float32)
        st = lib().or_avx_intersect_packets(self._h, ptr(rays), n, ptr(tri), ptr(dist), nthreads)
        if st:
            raise RuntimeError("oracle avx_intersect_packets status %d" % st)
        return tri, dist


def microfacet_normal_dist(h, n, gloss):
    return float(lib().or_microfacet_normal_dist(ptr(f32(h)), ptr(f32(n)), int(gloss)))


def microfacet_fresnel(incoming, h, color):
    out = np.zeros(3, dtype=np.float32)
    raw = np.zeros(1, dtype=np.float32)
    lib().or_microfacet_fresnel(ptr(f32(incoming)), ptr(f32(h)), ptr(f32(color)), ptr(out),
                                ptr(raw))
    return out, float(raw[0])


def microfacet_brdf(material, in_dir, out_dir, normal, ignore_fresnel=False,
                    approx=APPROX_PORTABLE):
    out = np.zeros(3, dtype=np.float32)
    raw = np.zeros(1, dtype=np.float32)
    lib().or_microfacet_brdf(int(material), ptr(f32(in_dir)), ptr(f32(out_dir)),
                             ptr(f32(normal)), 1 if ignore_fresnel else 0, approx, ptr(out),
                             ptr(raw))
    return out, float(raw[0])


def pd_brdf(normal, direction):
    return float(lib().or_pd_brdf(ptr(f32(normal)), ptr(f32(direction))))


def sky_intensity(direction):
    out = np.zeros(3, dtype=np.float32)
    lib().or_sky_intensity_lane(ptr(f32(direction)), ptr(out))
    return out


def avx_available():
    return bool(lib().or_avx_available())


def hdr_blocks_to_rows(width, height, hdr):
    rows = np.zeros((height, width, 3), dtype=np.float32)
    lib().or_hdr_blocks_to_rows(width, height, ptr(f32(hdr)), ptr(rows))
    return rows


def buffer_f32_into_bitmap(width, height, hdr, num_samples):
    bitmap = np.zeros(width * height, dtype=np.uint32)
    st = lib().or_buffer_f32_into_bitmap(width, height, ptr(f32(hdr)), ptr(bitmap), num_samples)
    if st:
        raise RuntimeError("status %d" % st)
    return bitmap
