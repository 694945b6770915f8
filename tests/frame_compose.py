"""Renderer::renderPath (frame 1, no sorting) composed from the fp32 oracle
stages on the CPU: the fp32 restatement of the whole frame, used to pin the
fp64 frame oracle and as the CPU side of the GPU stage-composition test."""
import numpy as np


def compose_frame_cpu(orc, hs, S2W, cfg, W, H):
    """Returns ([W*H, 4] float32 frame, per-bounce path-ray counts)."""
    from vulkanwavefrontpathtracer_amd import stages

    vec = lambda s: [float(x) for x in s.split()]  # noqa: E731
    cam, R = cfg["Camera"], cfg["Renderer"]
    light, bg = vec(cfg["Scene"]["light"]), vec(cfg["Scene"]["backgroundcolor"])
    S, radius, depth = R["numberOfShadowSamples"], R["shadowRadius"], R["recursionDepth"]
    i2p, _ = stages.pixel_table(W, H)
    npx = W * H
    tn = orc.tri_normals(hs.vertices, hs.indices)
    counts = []

    def trace(rays, closest):
        h = orc.trace_f32(rays, hs.tri_verts(), closest=closest)
        prim, geom = orc.hits_to_prim_geom(h[:, 0], hs.geom_first)
        o = h.copy()
        o[:, 0], o[:, 1], o[:, 2], o[:, 3] = prim, geom, h[:, 1], h[:, 2]
        return o

    s2i = i2p.copy()
    rays = orc.raygen_primary(S2W.T.reshape(-1), vec(cam["position"]), i2p, W, H, 0, cam["farPlane"])
    hits = trace(rays, True)
    seeds = orc.init_seeds(npx, 1)
    dec = np.ones((npx, 4), np.float32)
    fp = np.zeros((npx, 4), np.float32)
    aux = orc.reconstruct_smooth(hs.geom_first, tn, hs.tri_material, hs.emission, rays, hits, s2i,
                                 np.zeros((npx, 4), np.float32), dec, light, bg, 1)

    def shadow_c(rays, hits, s2i, aux, fp):
        n = len(rays)
        sr, ss, si = orc.raygen_shadow(seeds, rays, hits, s2i, 0, n, S, light, radius)
        return orc.reconstruct_shadow(trace(sr, False), s2i, si, aux, fp, 0, n, S, 0)

    fp = shadow_c(rays, hits, s2i, aux, fp)
    for _ in range(depth):
        out, os2i, seeds, dec = orc.shade_extend(hs.geom_first, tn, hs.tri_material, hs.albedo,
                                                 seeds, rays, hits, s2i, dec, False)
        counts.append(len(out))
        if len(out) == 0:
            break
        oh = trace(out, True)
        aux = orc.reconstruct_smooth(hs.geom_first, tn, hs.tri_material, hs.emission, out, oh, os2i,
                                     np.zeros((npx, 4), np.float32), dec, light, bg, 1)
        fp = shadow_c(out, oh, os2i, aux, fp)
        rays, hits, s2i = out, oh, os2i
    return fp, counts
