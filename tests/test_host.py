"""Host logic on the CPU: the C-ABI exports, env.json handling, the camera,
the primary-ray pixel order and the procedural scenes (no GPU calls)."""
import ctypes as C
import json
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_header_symbol():
    from vulkanwavefrontpathtracer_amd import _capi

    lib = _capi.load()
    header = (ROOT / "include" / "wfpt.h").read_text()
    declared = set(re.findall(r"\b(wf_[a-z0-9_]+)\s*\(", header))
    declared -= {"wf_scene_s", "wf_renderer_s"}
    missing = sorted(n for n in declared if getattr(lib, n, None) is None)
    assert not missing, f"declared in wfpt.h but not exported: {missing}"
    assert set(_capi.SIGNATURES) <= declared | {"wf_last_error"}
    assert _capi.missing_symbols() == []


def test_status_strings():
    from vulkanwavefrontpathtracer_amd import _capi

    lib = _capi.load()
    assert lib.wf_status_string(0) == b"ok"
    assert lib.wf_status_string(-4) == b"scratch buffer too small"
    assert b"gfx950" in lib.wf_version()


def test_sort_scratch_sizes():
    from vulkanwavefrontpathtracer_amd import _capi

    lib = _capi.load()
    a = lib.wf_sort_scratch_bytes(1 << 20, 4, 32, 8)
    b = lib.wf_sort_scratch_bytes(1 << 21, 4, 32, 8)
    assert 0 < a < b
    assert lib.wf_sort_scratch_bytes(10, 4, 32, 5) < 0        # bad digit width
    assert lib.wf_sort_scratch_bytes(10, 4, 40, 8) < 0        # more bits than the key holds
    assert lib.wf_sort_scratch_bytes(10, 8, 64, 4) > 0
    assert lib.wf_sort_scratch_bytes(0, 8, 64, 8) > 0


# ----------------------------------------------------------------- env.json
def renderer_from(obj):
    from vulkanwavefrontpathtracer_amd.renderer import Renderer

    return Renderer(obj)


def test_env_defaults_match_app_environment():
    r = renderer_from({})
    # AppEnvironment.cpp:12-41
    assert r.env_string("Application.mode") == "interactive"
    assert r.env_int("Resolution.width") == 1024 and r.env_int("Resolution.height") == 768
    assert r.env_vector("Scene.light") == [5.0, -2.0, 1.0]
    assert r.env_int("Renderer.numberOfShadowSamples") == 2
    assert r.env_float("Renderer.shadowRadius") == pytest.approx(0.001)
    assert r.env_string("Renderer.rayType") == "primary"
    assert r.env_float("Camera.fieldOfView") == pytest.approx(45.0)
    assert r.size == (1024, 768)


def test_env_nested_override_and_types():
    r = renderer_from({
        "Resolution": {"width": 320, "height": 200},
        "Renderer": {"rayType": "path", "recursionDepth": 4, "sortPathRays": True,
                     "keyValue": 0.5},
        "Scene": {"light": "1 2 3", "unknown": 7},
        "Unknown": {"x": 1},
    })
    assert r.size == (320, 200)
    assert r.env_string("Renderer.rayType") == "path"
    assert r.env_int("Renderer.recursionDepth") == 4
    assert r.env_int("Renderer.sortPathRays") == 1
    assert r.env_float("Renderer.keyValue") == pytest.approx(0.5)
    assert r.env_vector("Scene.light") == [1.0, 2.0, 3.0]


def test_env_reference_file_format(tmp_path):
    # the shape of the reference's env/env.json (Application/Benchmark/...)
    text = {
        "Application": {"mode": "benchmark"}, "Benchmark": {"printSortLogs": True},
        "Resolution": {"width": 64, "height": 32},
        "Scene": {"filename": "models/x.gltf", "light": "0.0 10.0 0.0",
                  "backgroundcolor": "0.529 0.808 0.98", "headlight": True},
        "Renderer": {"numberOfPrimarySamples": 1, "numberOfShadowSamples": 2,
                     "shadowRadius": 1.0, "rayType": "path", "recursionDepth": 4},
        "Camera": {"position": "8.287917 3.804885 34.659367",
                   "rotation": "1.075014 -303.599152 0.000000", "nearPlane": 0.1,
                   "farPlane": 512.0, "fieldOfView": 60.0},
    }
    p = tmp_path / "env.json"
    p.write_text(json.dumps(text))
    from vulkanwavefrontpathtracer_amd.renderer import Renderer

    r = Renderer(p)
    assert r.env_string("Scene.filename") == "models/x.gltf"
    assert r.env_float("Camera.farPlane") == pytest.approx(512.0)
    assert r.env_vector("Camera.position") == pytest.approx([8.287917, 3.804885, 34.659367])


def test_env_errors(tmp_path):
    from vulkanwavefrontpathtracer_amd import _capi
    from vulkanwavefrontpathtracer_amd.renderer import Renderer

    p = tmp_path / "bad.json"
    p.write_text('{"Resolution": {"width": 3,, }')
    with pytest.raises(_capi.WfError) as e:
        Renderer(p)
    assert e.value.status == -6
    with pytest.raises(_capi.WfError) as e:
        Renderer(tmp_path / "missing.json")
    assert e.value.status == -7
    with pytest.raises(_capi.WfError):
        Renderer({"Resolution": {"width": 0, "height": 10}})
    with pytest.raises(_capi.WfError):
        Renderer({"Renderer": {"keyMode": "nonsense"}})


def test_bundled_configs_parse():
    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    sizes = {"config0.json": (256, 256), "config1.json": (1920, 1080),
             "config2.json": (1920, 1080), "config4.json": (512, 512)}
    for name, size in sizes.items():
        assert Renderer(env_path(name)).size == size


# ------------------------------------------------------------------- camera
def test_camera_forward_and_inverse():
    from vulkanwavefrontpathtracer_amd import stages

    V, P, S = stages.camera_matrices([1, 2, 3], [0, 0, 0], 60.0, 1.5, 0.1, 100.0)
    assert np.allclose(S @ (P @ V), np.eye(4), atol=1e-4)
    # glm::perspective RH zero-to-one with Y flipped (camera.hpp:108-120)
    th = np.tan(np.radians(30.0))
    assert P[0, 0] == pytest.approx(1 / (1.5 * th), rel=1e-6)
    assert P[1, 1] == pytest.approx(-1 / th, rel=1e-6)
    assert P[2, 2] == pytest.approx(100.0 / (0.1 - 100.0), rel=1e-6)
    assert P[3, 2] == -1.0
    # the view translates the camera to the origin
    assert np.allclose(V @ np.array([1, 2, 3, 1.0]), [0, 0, 0, 1], atol=1e-6)


@pytest.mark.parametrize("rot,forward", [((0, 0, 0), (0, 0, -1)), ((0, 90, 0), (-1, 0, 0)),
                                         ((0, -90, 0), (1, 0, 0)), ((30, 0, 0), (0, 0.5, -0.866025))])
def test_camera_rotation_conventions(rot, forward):
    from vulkanwavefrontpathtracer_amd import stages

    V, P, S = stages.camera_matrices([0, 0, 0], rot, 60.0, 1.0, 0.1, 100.0)
    # the world-space view direction is -z of the camera frame = -(third row of V)
    f = -V[2, :3]
    assert np.allclose(f, forward, atol=1e-5)


# ------------------------------------------------------------- pixel table
def test_pixel_table_small_known_answer():
    from vulkanwavefrontpathtracer_amd import stages

    i2p, p2i = stages.pixel_table(16, 8)
    # first block (8x8 at the origin) is Z-swizzled: inner 0..3 -> (0,0) (1,0) (0,1) (1,1)
    assert list(i2p[:4]) == [0, 1, 16, 17]
    assert list(i2p[4:8]) == [2, 3, 18, 19]
    # the second 8x8 block (x in 8..15) follows the first
    assert i2p[64] == 8
    assert np.array_equal(np.sort(i2p), np.arange(128))
    assert np.array_equal(p2i[i2p], np.arange(128))


def _pixel_order_by_sort(w, h):
    """The slot order as a lexicographic sort of every pixel by (region,
    primary, secondary): region 0 is the 8-aligned part, ordered by the
    Morton code of its 8x8 block and then by the Morton code inside the
    block; region 1 the bottom stripe ordered by (x, y); region 2 the right
    stripe with the corner ordered by (y, x)."""
    def morton(x, y):
        code = np.zeros_like(x)
        for bit in range(16):
            code |= ((x >> bit) & 1) << (2 * bit) | ((y >> bit) & 1) << (2 * bit + 1)
        return code

    y, x = np.divmod(np.arange(w * h, dtype=np.int64), w)
    bw, bh = w & ~7, h & ~7
    region = np.where((x < bw) & (y < bh), 0, np.where(x < bw, 1, 2))
    primary = np.select([region == 0, region == 1], [morton(x >> 3, y >> 3), x], y)
    secondary = np.select([region == 0, region == 1], [morton(x & 7, y & 7), y], x)
    return np.lexsort((secondary, primary, region))


@pytest.mark.parametrize("w,h", [(1, 1), (8, 8), (16, 8), (7, 5), (13, 21), (100, 37), (256, 256),
                                 (203, 117), (1920, 1080), (72, 1000)])
def test_pixel_table_equals_sorted_definition(w, h):
    """Every slot of the table equals the ordering's definition restated as a
    numpy sort (sizes with and without stripes, wide and tall images whose
    block grid is not a power of two)."""
    from vulkanwavefrontpathtracer_amd import stages

    i2p, p2i = stages.pixel_table(w, h)
    exp = _pixel_order_by_sort(w, h)
    assert np.array_equal(i2p, exp)
    assert np.array_equal(p2i[exp], np.arange(w * h))


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (13, 21), (256, 256), (1920, 1080), (100, 37)])
def test_pixel_table_bijection_with_stripes(w, h):
    from vulkanwavefrontpathtracer_amd import stages

    i2p, p2i = stages.pixel_table(w, h)
    assert np.array_equal(np.sort(i2p), np.arange(w * h))
    assert np.array_equal(p2i[i2p], np.arange(w * h))
    bw, bh = w & ~7, h & ~7
    nb = bw * bh
    # bulk first: every pixel inside the 8-aligned block region comes before the stripes
    xs, ys = i2p % w, i2p // w
    assert np.all(xs[:nb] < bw) and np.all(ys[:nb] < bh)
    if nb < w * h:
        assert np.all((xs[nb:] >= bw) | (ys[nb:] >= bh))


# --------------------------------------------------------- procedural scenes
def test_boxes_scene():
    from vulkanwavefrontpathtracer_amd import stages

    s = stages.HostScene(stages.SCENE_BOXES, 1000)
    assert s.num_tris == 64 * 192 + 2 + 2 == 12292
    assert len(s.geom_first) == 64 + 2 + 1 + 0 and s.geom_first[-1] == s.num_tris
    assert np.all(np.diff(s.geom_first) > 0)
    assert s.albedo.shape == (66, 3)
    assert np.all((s.albedo[:64] >= 0.1) & (s.albedo[:64] <= 0.9))
    assert np.allclose(s.emission[65], 10.0)
    # unit normals, and bounds contain every vertex
    n = np.linalg.norm(s.vertices[:, 3:], axis=1)
    assert np.allclose(n, 1.0, atol=1e-5)
    v = s.vertices[:, :3]
    assert np.all(v >= s.bmin - 1e-6) and np.all(v <= s.bmax + 1e-6)
    s2 = stages.HostScene(stages.SCENE_BOXES, 1000)
    assert np.array_equal(s.vertices, s2.vertices)          # seeded, deterministic
    s3 = stages.HostScene(stages.SCENE_BOXES, 1001)
    assert not np.array_equal(s.vertices, s3.vertices)


def test_interior_scene_size():
    from vulkanwavefrontpathtracer_amd import stages

    s = stages.HostScene(stages.SCENE_INTERIOR, 1001)
    assert abs(s.num_tris - 262_144) / 262_144 < 0.01
    assert s.num_tris == 12 + 32 * 4096 + 24 * 3840 + 40_000 + 6
    assert s.albedo.shape[0] == 9 and np.allclose(s.emission[8], 10.0)
    assert int(s.tri_material.max()) <= 8 and int(s.tri_material.min()) >= 0


def test_material_groups_regroup_interior():
    """Scene.materialGroups: emissive materials first, then contiguous runs of
    the other triangles, each inheriting its first triangle's albedo."""
    from vulkanwavefrontpathtracer_amd import stages

    base = stages.HostScene(stages.SCENE_INTERIOR, 1001)
    g = stages.HostScene(stages.SCENE_INTERIOR, 1001, material_groups=512)
    assert g.albedo.shape == (512, 3) and np.array_equal(g.indices, base.indices)
    emissive = np.any(base.emission[base.tri_material] > 0, axis=1)
    ne = int(np.any(base.emission > 0, axis=1).sum())
    assert np.all(g.tri_material[emissive] < ne) and np.all(g.tri_material[~emissive] >= ne)
    plain = g.tri_material[~emissive]
    assert np.all(np.diff(plain) >= 0)                       # contiguous runs in order
    assert len(np.unique(plain)) == 512 - ne
    first = np.searchsorted(plain, np.arange(ne, 512))
    src = base.tri_material[~emissive][first]
    assert np.array_equal(g.albedo[ne:], base.albedo[src])
    assert np.all(g.emission[ne:] == 0)


def test_adjoint_argument_checks_without_gpu():
    from vulkanwavefrontpathtracer_amd import _capi
    from vulkanwavefrontpathtracer_amd.renderer import Renderer, env_path

    r = Renderer(env_path("config0.json"))
    with pytest.raises(_capi.WfError, match="null"):
        r_call = _capi.call
        r_call("wf_renderer_adjoint", r.handle, 1, None, None, None, None, None)
    with pytest.raises(_capi.WfError, match="no scene"):
        _capi.call("wf_renderer_adjoint", r.handle, 1, 64, 64, 64, None, None)


def test_stage_entry_points_check_arguments_without_gpu():
    """Every stage entry point of the C ABI, the way the reference's stage
    calls get their counts: n = 0 (an empty wavefront) returns WF_OK and
    touches nothing, n < 0 and null buffers with n > 0 return
    WF_ERR_INVALID_ARG.  All of it is decided before any device call, so it
    runs on the CPU (buffer and scene arguments are never dereferenced on
    these paths; a non-null stand-in marks "present")."""
    from vulkanwavefrontpathtracer_amd import _capi

    lib = _capi.load()
    OK, BAD = 0, -1
    P = C.c_void_p(0x1000)          # present (never dereferenced on these paths)
    N = None                        # absent
    f3 = (C.c_float * 3)(0.0, 0.0, 0.0)
    calls = {
        "wf_morton_keys": lambda n, b: lib.wf_morton_keys(b, b, b, n, 0, f3, f3, None, None),
        "wf_sort_pairs": lambda n, b: lib.wf_sort_pairs(b, b, n, 4, 32, 8, b, 1 << 40, None, None),
        "wf_reorder": lambda n, b: lib.wf_reorder(b, b, P, b, b, b, n, None, None),
        "wf_trace": lambda n, b: lib.wf_trace(P, b, None, b, n, 1, None, None),
        "wf_reorder_trace": lambda n, b: lib.wf_reorder_trace(P, b, b, P, b, b, b, b, n, 1, None, None),
        "wf_trace_stats": lambda n, b: lib.wf_trace_stats(P, b, b, n, 1, b, b, None),
        "wf_init_seeds": lambda n, b: lib.wf_init_seeds(b, n, 1, None, None),
        "wf_fill_f32": lambda n, b: lib.wf_fill_f32(b, n, 1.0, None, None),
        "wf_count_hits": lambda n, b: lib.wf_count_hits(b, n, P, None, None),
        "wf_interpolate_colors": lambda n, b: lib.wf_interpolate_colors(b, b, n, 1, 0.6, 2.0, None, None),
        "wf_reconstruct_smooth": lambda n, b: lib.wf_reconstruct_smooth(P, b, b, b, b, b, n, f3, f3, 1,
                                                                          None, None),
        "wf_reconstruct_shadow": lambda n, b: lib.wf_reconstruct_shadow(b, b, b, b, b, 0, n, 2, 0,
                                                                          None, None),
        "wf_raygen_shadow": lambda n, b: lib.wf_raygen_shadow(b, b, b, b, b, b, b, 0, n, 2, f3, 1.0,
                                                                None, None),
    }
    for name, fn in calls.items():
        assert fn(0, N) == OK, f"{name}: empty input"
        assert fn(-1, P) == BAD, f"{name}: negative count"
        assert fn(5, N) == BAD, f"{name}: null buffers"
    # in-place reorders are refused
    assert lib.wf_reorder(P, P, P, P, P, P, 4, None, None) == BAD
    assert lib.wf_reorder_trace(P, P, P, P, P, P, P, P, 4, 1, None, None) == BAD
    # unknown key modes and digit widths
    assert lib.wf_morton_keys(P, P, P, 4, 7, f3, f3, None, None) == BAD
    assert lib.wf_sort_pairs(P, P, 4, 4, 32, 5, P, 1 << 40, None, None) == BAD
    assert lib.wf_sort_pairs(P, P, 4, 4, 33, 8, P, 1 << 40, None, None) == BAD
    # a sort scratch smaller than wf_sort_scratch_bytes
    need = lib.wf_sort_scratch_bytes(1000, 4, 32, 8)
    assert lib.wf_sort_pairs(P, P, 1000, 4, 32, 8, P, need - 1, None, None) != OK


def test_sized_boxes_and_batched_workload_spec():
    from vulkanwavefrontpathtracer_amd import stages
    from vulkanwavefrontpathtracer_amd.batch import workload

    hs = stages.HostScene(0, 3000, num_boxes=10)
    assert hs.num_tris == 10 * 192 + 4 and hs.albedo.shape == (12, 3)
    assert hs.emission[11].max() == 10.0                      # the light quad's material
    base = stages.HostScene(0, 1000)
    same = stages.HostScene(0, 1000, num_boxes=64)
    assert np.array_equal(base.vertices, same.vertices)        # 64 boxes = the boxes scene
    boxes, seeds, views = workload()
    tris = np.array(boxes) * 192
    assert len(views) == 64 and seeds[0] == 3000 and seeds[-1] == 3063
    assert tris.min() >= 900 and tris.max() <= 500_200
    assert {v["width"] for v in views} <= {128, 256, 512, 1024}
    assert all(1 <= v["spp"] <= 16 for v in views)
    assert np.log(tris).std() > 1.0                            # spread over decades


def test_png_and_pfm_writers(tmp_path):
    """present.comp's rgba8 conversion (clamp, x * 255 to nearest) in a PNG
    that zlib and the PNG chunk CRCs accept; PFM rows bottom to top."""
    import struct
    import zlib

    from vulkanwavefrontpathtracer_amd import _capi

    w, h = 70, 3                                    # > 65535-byte stored blocks need not apply
    rng = np.random.default_rng(2)
    img = rng.uniform(-0.2, 1.2, (h * w, 4)).astype(np.float32)
    img[0, :3] = [np.nan, 0.5 / 255, 1.0]
    p = tmp_path / "a.png"
    _capi.call("wf_write_png", str(p).encode(), img.ctypes.data_as(_capi.pf), w, h)
    data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body)
        chunks[typ] = chunks.get(typ, b"") + body
        pos += 12 + n
    assert struct.unpack(">IIBBBBB", chunks[b"IHDR"]) == (w, h, 8, 2, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(h, 3 * w + 1)
    assert np.all(raw[:, 0] == 0)
    got = raw[:, 1:].reshape(h * w, 3)
    x = np.nan_to_num(img[:, :3], nan=0.0)
    want = np.where(x <= 0, 0, np.where(x >= 1, 255, np.rint(x * 255))).astype(np.uint8)
    assert np.array_equal(got, want)
    q = tmp_path / "a.pfm"
    _capi.call("wf_write_pfm", str(q).encode(), img.ctypes.data_as(_capi.pf), w, h)
    head, rest = q.read_bytes().split(b"\n-1.0\n", 1)
    assert head == f"PF\n{w} {h}".encode()
    pf = np.frombuffer(rest, "<f4").reshape(h, w, 3)[::-1].reshape(h * w, 3)
    assert np.array_equal(pf, img[:, :3], equal_nan=True)


# ------------------------------------------------------- C ABI as compiled code
def _integration_example():
    """The C++ frame loop of INTEGRATION.md section 3, as written there."""
    import re

    text = (ROOT / "INTEGRATION.md").read_text()
    sec = text[text.index("## 3."):text.index("## 4.")]
    return re.search(r"```cpp\n(.*?)```", sec, re.S).group(1)


def build_integration_binary(tmp_path):
    """g++ on the documented example plus a main(argv[1] = env path) that
    prints render_one_frame's status, linked against libwfpt.so."""
    import subprocess

    from vulkanwavefrontpathtracer_amd import _capi

    lib = Path(_capi.load()._name)
    src = tmp_path / "frame_loop.cpp"
    src.write_text(_integration_example() + """
#include <cstdio>
int main(int argc, char** argv) {
    if (argc < 2) {               // link check only: host-side calls, no device
        wf_renderer r = nullptr;
        int s = wf_renderer_create_from_string("{\\"Resolution\\": {\\"width\\": 16}}", &r);
        int w = 0;
        s |= wf_env_get_int(r, "Resolution.width", &w);
        wf_renderer_destroy(r);
        std::printf("host %d %d %s\\n", s, w, wf_version());
        return s;
    }
    int s = render_one_frame(argv[1]);
    std::printf("frame %d\\n", s);
    return s;
}
""")
    exe = tmp_path / "frame_loop"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
           str(src), "-o", str(exe), f"-L{lib.parent}", "-lwfpt", f"-Wl,-rpath,{lib.parent}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_integration_example_compiles_links_and_runs(tmp_path):
    """The reference's C++ host binds by #include and link: the documented
    example builds against wfpt.h and libwfpt.so and its host-side calls run."""
    import subprocess

    exe = build_integration_binary(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("host 0 16 wfpt-mi355x")


def test_header_compiles_as_c99(tmp_path):
    """wfpt.h is a C header: every declaration compiles as strict C99."""
    import subprocess

    src = tmp_path / "use.c"
    src.write_text("""#include "wfpt.h"
int probe(wf_scene scene, const wf_ray* rays, wf_hit* hits, int64_t n, wf_stream_t s) {
    wf_ray_buffer rb = {0};
    wf_sort_log log = {0};
    (void)rb; (void)log;
    return wf_trace(scene, rays, 0, hits, n, 1, s, 0) == WF_OK ? wf_key_bytes(WF_KEY_AILA32) : -1;
}
""")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", f"-I{ROOT / 'include'}",
           "-c", str(src), "-o", str(tmp_path / "use.o")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
