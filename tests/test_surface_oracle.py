"""Textured and alpha-tested surfaces in the oracle (surfaceAt, src/core/voxel.ts:325-336; the base-colour array and
its mips, src/core/scene_gpu.ts:38-91; shadeSample, src/core/dynlayer.ts:451-452), pinned by what the mathematics
fixes: the sRGB transfer function of IEC 61966-2-1 in float64, exact round trips, float64 box means and bilinear
interpolation, repeat addressing, and hand-worked alpha-test and albedo closed forms on one textured quad.

The reference samples through Dawn's texture units, whose filtering precision is the driver's; it cannot run here
(no Dawn, no Vulkan ICD), so the sampler is pinned against these closed forms, not against the reference's bits."""
import dataclasses

import numpy as np

import rtgi
from conftest import textured
from oracle import oracle as orc
from oracle.oracle import DynState, Oracle, SurfaceData


def srgb_decode64(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def srgb_encode64(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.maximum(x, 0) ** (1 / 2.4) - 0.055)


def _surface(images, layer=(1,), flags=(0,), cutoff=(0.5,), nverts=1):
    return rtgi.Surface(uvs=np.zeros((nverts, 2), np.float32), images=np.asarray(images, np.uint8),
                        layer=np.asarray(layer, np.uint32), flags=np.asarray(flags, np.uint32),
                        cutoff=np.asarray(cutoff, np.float32))


def test_srgb_tables_are_the_transfer_function():
    lut = orc.srgb_tables()
    np.testing.assert_array_equal(lut[:256], srgb_decode64(np.arange(256) / 255.0).astype(np.float32))
    np.testing.assert_array_equal(lut[256:], srgb_decode64((np.arange(255) + 0.5) / 255.0).astype(np.float32))
    assert lut[0] == 0.0 and lut[255] == 1.0 and lut[10] == np.float32(10 / 255 / 12.92)
    assert np.all(np.diff(lut[:256]) > 0) and np.all(np.diff(lut[256:]) > 0)
    # each threshold lies between the decoded values of its two codes
    assert np.all(lut[256:] > lut[:255]) and np.all(lut[256:] < lut[1:256])


def test_mips_keep_constant_layers():
    """Every code, as a constant 2 x 2 layer, comes back unchanged from its 1 x 1 mip (encode(decode(c)) = c, alpha
    too): the mip of equal texels is that texel."""
    img = np.zeros((256, 2, 2, 4), np.uint8)
    img[..., :3] = np.arange(256)[:, None, None, None]
    img[..., 3] = np.arange(256)[:, None, None]
    s = SurfaceData(_surface(img))
    top = s.tex.reshape(256, 5)[:, 4]
    want = np.arange(256, dtype=np.uint32) * 0x01010101
    np.testing.assert_array_equal(top, want)


def test_mips_are_box_means_in_linear_space():
    """Level m texel = the mean of its 2 x 2 parents decoded from sRGB, stored as the nearest code of the sRGB
    encoding (float64 reference; a code may differ only where the mean is within 1e-6 of a rounding boundary)."""
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (3, 16, 16, 4)).astype(np.uint8)
    s = SurfaceData(_surface(img, layer=(1, 2, 3), flags=(0, 0, 0), cutoff=(0, 0, 0)))
    per = s.per
    levels = [img.astype(np.int64)]
    off = 0
    for m in range(1, 5):
        off += (16 >> (m - 1)) ** 2
        w = 16 >> m
        prev = levels[-1]
        dec = srgb_decode64(prev[..., :3] / 255.0)
        mean = dec.reshape(3, w, 2, w, 2, 3).mean(axis=(2, 4))
        code = np.clip(np.floor(255 * srgb_encode64(mean) + 0.5), 0, 255).astype(np.int64)
        asum = prev[..., 3].reshape(3, w, 2, w, 2).sum(axis=(2, 4))
        acode = (asum * 2 + 4) // 8  # floor(sum / 4 + 0.5)
        got = s.tex.reshape(3, per)[:, off:off + w * w].reshape(3, w, w)
        grgb = np.stack([(got >> (8 * k)) & 255 for k in range(3)], -1).astype(np.int64)
        ga = (got >> 24).astype(np.int64)
        bounds = srgb_decode64((np.arange(255) + 0.5) / 255.0)
        near = np.min(np.abs(mean[..., None] - bounds), axis=-1) < 1e-6
        assert np.all((grgb == code) | near), m
        assert np.all((ga == acode) | (asum % 4 == 2)), m
        levels.append(np.concatenate([grgb, ga[..., None]], -1))


def test_bilinear_sample_repeat_and_levels():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (2, 32, 32, 4)).astype(np.uint8)
    s = SurfaceData(_surface(img, layer=(1, 2), flags=(0, 0), cutoff=(0, 0)))
    lut = s.lut.astype(np.float64)

    def level(layer, m):
        w = 32 >> m
        off = sum((32 >> j) ** 2 for j in range(m))
        t = s.tex.reshape(2, s.per)[layer - 1, off:off + w * w].reshape(w, w)
        return np.stack([lut[t & 255], lut[(t >> 8) & 255], lut[(t >> 16) & 255], (t >> 24) / 255.0], -1)

    for lod in range(7):
        m = min(lod, 5)
        w = 32 >> m
        T = level(2, m)
        for _ in range(20):
            i, j = rng.integers(0, w, 2)
            # texel centres return the texel (weights exactly 1 and 0)
            np.testing.assert_array_equal(s.sample(2, (i + 0.5) / w, (j + 0.5) / w, lod), T[j, i].astype(np.float32))
            # any point: float64 bilinear of the four neighbours, repeat-wrapped, at the f32 texel coordinate
            u, v = rng.uniform(-3, 3, 2).astype(np.float32)
            x = float(np.float32(u * np.float32(w)) - np.float32(0.5))
            y = float(np.float32(v * np.float32(w)) - np.float32(0.5))
            x0, y0 = int(np.floor(x)), int(np.floor(y))
            fx, fy = x - x0, y - y0
            q = lambda a, b: T[b % w, a % w]  # noqa: E731
            want = ((q(x0, y0) * (1 - fx) + q(x0 + 1, y0) * fx) * (1 - fy) +
                    (q(x0, y0 + 1) * (1 - fx) + q(x0 + 1, y0 + 1) * fx) * fy)
            np.testing.assert_allclose(s.sample(2, u, v, lod), want, atol=1e-6)
            # repeat addressing: whole-number shifts of dyadic coordinates sample the same point
            ud, vd = np.float32(i / 64.0 + 1 / 256.0), np.float32(j / 64.0)
            np.testing.assert_array_equal(s.sample(2, ud, vd, lod), s.sample(2, ud + 3.0, vd - 2.0, lod))
    assert np.array_equal(s.sample(0, 0.3, 0.7, 3), np.ones(4, np.float32))  # layer 0: the white default
    # a non-finite texel coordinate reads texel 0 with weight 1
    assert np.array_equal(s.sample(1, np.inf, np.nan, 0), s.sample(1, 0.5 / 32, 0.5 / 32, 0))


def floor_world(flags, layer_img, factor=(0.5, 1.0, 1.0, 1.0), cutoff=0.5):
    """One 7.9 x 3.9 m quad in the plane y = 0.1 (inside the first voxel row of v0 = 0.25) with u = x / 8, v = z / 8,
    one material sampling layer 1 of a 64 x 64 texture."""
    pos = np.array([[0.05, 0.1, 0.05], [7.95, 0.1, 0.05], [7.95, 0.1, 3.95], [0.05, 0.1, 3.95]], np.float32)
    nrm = np.tile(np.array([[0.0, 1.0, 0.0]], np.float32), (4, 1))
    tri = np.array([[0, 1, 2, 0], [0, 2, 3, 0]], np.uint32)
    mat = np.array([factor], np.float32)
    surf = rtgi.Surface(uvs=pos[:, [0, 2]] / 8.0, images=layer_img[None], layer=np.array([1], np.uint32),
                        flags=np.array([flags], np.uint32), cutoff=np.array([cutoff], np.float32))
    scene = rtgi.Scene(pos=pos, nrm=nrm, tri=tri, materials=mat, surface=surf)
    vox = rtgi.VoxParams(v0=0.25, levels=2, G=8, maxBricks=2 * 512, maxVoxels=1 << 16)
    rc = rtgi.RcParams(cascades=2, probes0=8, s0=1.0)
    return scene, vox, rc, rtgi.Frame(eye=(4.0, 0.5, 2.0))


def _cells(cm, E):
    occ = np.unpackbits(cm["occ"].view(np.uint8), bitorder="little")[:E ** 3].astype(bool)
    return occ


def test_alpha_test_and_albedo_closed_form():
    """Left half of the texture alpha 0, right half 255, rgb codes (200, 100, 50): at level 0 (mip 3, 8 x 8 texels)
    the alpha at a voxel centre x is the linear ramp between texel columns 3 and 4, so alpha >= 0.5 exactly for the
    voxels with x > 4 m (and by the repeat wrap for none on the left: x = 0.125 m samples columns 7 and 0 at
    weights 0.375 / 0.625). The masked build keeps exactly those voxels of the unmasked one; every kept voxel's
    albedo is decode(code) x factor with the toon flag in alpha."""
    img = np.zeros((64, 64, 4), np.uint8)
    img[..., :3] = (200, 100, 50)
    img[:, 32:, 3] = 255
    scene, vox, rc, frame = floor_world(rtgi.SURF_ALPHA_MASK | rtgi.SURF_TOON, img)
    E = vox.G * 8
    org = orc.origins(vox.v0, 1, vox.G, frame.eye)[0] * 8
    masked = Oracle(vox, rc, frame)
    masked.build(scene)
    plain = Oracle(vox, rc, frame)
    plain.build(dataclasses.replace(scene, surface=dataclasses.replace(scene.surface, flags=np.array([0], np.uint32))))
    cm, cp = masked.clipmap(0), plain.clipmap(0)
    om, op = _cells(cm, E), _cells(cp, E)
    idx = np.arange(E ** 3)
    xc = (org[0] + idx % E + 0.5) * vox.v0
    assert op.sum() == 32 * 16 and np.array_equal(om, op & (xc > 4.0))
    lut = orc.srgb_tables()
    want = orc.pack4x8unorm([lut[200] * 0.5, lut[100], lut[50], 1.0])
    assert np.all(cm["albedo"][om] == want)
    # the unmasked, non-toon build samples the same colour (alpha plays no part) and keeps every voxel
    assert np.all(cp["albedo"][op] == orc.pack4x8unorm([lut[200] * 0.5, lut[100], lut[50], 0.0]))


def test_untextured_equals_white_layer():
    """A surface whose materials all sample layer 0 (the white default) and are unmasked builds the untextured
    clipmap bit for bit (white x factor = factor), also on a generated atrium."""
    scene, vox, rc, frame = rtgi.config("atrium")
    nm = scene.materials.shape[0]
    surf = rtgi.Surface(uvs=np.zeros((scene.pos.shape[0], 2), np.float32), images=np.zeros((1, 4, 4, 4), np.uint8),
                        layer=np.zeros(nm, np.uint32), flags=np.zeros(nm, np.uint32), cutoff=np.zeros(nm, np.float32))
    a = Oracle(vox, rc, frame)
    a.build(scene)
    b = Oracle(vox, rc, frame)
    b.build(dataclasses.replace(scene, surface=surf))
    assert np.array_equal(a.buf, b.buf) and np.array_equal(a.rep, b.rep)


def test_textured_atrium_masks_and_recolours():
    """On the textured atrium, masked materials lose voxels and nothing else does: at every level each voxel of the
    textured build is a voxel of the untextured one, voxels are lost (the masked half's transparent bar), and every
    lost voxel's representative triangle (the highest of those touching it) has a masked material."""
    scene, vox, rc, frame = rtgi.config("atrium")
    ts = textured(scene)
    a = Oracle(vox, rc, frame)
    a.build(scene)
    b = Oracle(vox, rc, frame)
    b.build(ts)
    E = vox.G * 8
    lost_total = 0
    for L in range(vox.levels):
        ca, cb = a.clipmap(L), b.clipmap(L)
        oa, ob = _cells(ca, E), _cells(cb, E)
        assert np.all(oa | ~ob)
        lost = oa & ~ob
        lost_total += lost.sum()
        rep = ca["rep"][lost].astype(np.int64) - 1
        assert np.all(ts.surface.flags[scene.tri[rep, 3]] & rtgi.SURF_ALPHA_MASK)
        assert np.any(ca["albedo"][ob] != cb["albedo"][ob])
    assert lost_total > 1000


def test_dynamic_layer_drops_transparent_samples():
    """A dynamic square of a masked material with alpha 0 everywhere splats nothing (every sample fails the test:
    the layer's words stay 0), while its bricks are still marked and allocated (marking does not sample); with
    alpha 255 it splats what the untextured square does, up to the sampled albedo."""
    from test_oracle import dyn_world
    scene, vox, rc, frame = dyn_world()
    img = np.zeros((1, 16, 16, 4), np.uint8)
    img[..., :3] = 255
    nm = scene.materials.shape[0]
    surf = rtgi.Surface(uvs=np.zeros((scene.pos.shape[0], 2), np.float32), images=img,
                        layer=np.ones(nm, np.uint32), flags=np.full(nm, rtgi.SURF_ALPHA_MASK, np.uint32),
                        cutoff=np.full(nm, 0.5, np.float32))
    ref = DynState(vox, 64, scene.firstDynamicTri)
    Oracle(vox, rc, frame).build(scene, dyn=ref)
    clear = DynState(vox, 64, scene.firstDynamicTri)
    Oracle(vox, rc, frame).build(dataclasses.replace(scene, surface=surf), dyn=clear)
    assert clear.hdr[8] == ref.hdr[8] and np.array_equal(clear.ptr, ref.ptr)
    assert not np.any(clear.data[:clear.hdr[8]])
    img[..., 3] = 255
    opaque = DynState(vox, 64, scene.firstDynamicTri)
    Oracle(vox, rc, frame).build(dataclasses.replace(scene, surface=surf), dyn=opaque)
    n = ref.hdr[8]
    np.testing.assert_array_equal(opaque.data[:n, :, :7], ref.data[:n, :, :7])  # areas: every sample kept
