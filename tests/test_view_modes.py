"""View modes of renderToBuffer (treeletRenderer_rgba.cl:644-732) in the
oracle, pinned on the CPU against independent numpy restatements of each
mode's formula applied to the oracle's own per-ray outputs (hit node, t,
depth, loop count, centre) and the scene's attributes."""
import numpy as np
import pytest

import libsvo_amd as svo
from libsvo_amd import workloads

F = np.float32
K = F(0.003921569)


def _unpack(v):
    v = np.asarray(v, np.uint32)
    return np.stack([(v >> 24) & 255, (v >> 16) & 255, (v >> 8) & 255, v & 255], -1).astype(np.float32)


def _render(menger8, mode, lod=0.0, seed=7):
    import oracle
    sc, base = menger8
    ora = oracle.OracleScene(base, workloads.incore_bytes_for(sc))
    cam = svo.with_view_mode(workloads.sphere_camera(seed, 96, 80, lod=lod), mode)
    r = ora.render(cam, "treelet", shade=True)
    inc = ora.incore()
    return cam, r, inc


def test_color_normal_depth_t_loops(menger8):
    for mode in ("color", "normal", "svo_depth", "t", "loops"):
        cam, r, inc = _render(menger8, mode)
        hit = r["node"] != svo.MISS
        s = r["shade"]
        assert np.all(s[..., 3] == 1.0)
        if mode == "color":
            want = _unpack(inc["colors"][r["node"][hit]])[:, :3] * K
            assert np.array_equal(s[hit][:, :3], want)
            assert np.all(s[~hit][:, :3] == np.array([0.2, 0.3, 0.2], np.float32))
        elif mode == "normal":
            want = (_unpack(inc["normals"][r["node"][hit]])[:, :3] * K - F(0.5)) * F(2.0)
            assert np.array_equal(s[hit][:, :3], want)
        elif mode == "svo_depth":
            want = r["depth"].astype(np.float32) / F(13) * hit.astype(np.float32)
            assert np.array_equal(s[..., 0], want)
        elif mode == "t":
            assert np.array_equal(s[..., 0], F(1.0) - r["t"] * F(0.1))
        else:
            assert np.array_equal(s[..., 0], F(1.0) - r["loops"].astype(np.float32) / F(200.0))
        assert hit.sum() > 1000 and (~hit).sum() > 100


def test_quality_is_the_fourth_root_of_the_leaf_size_without_lod(menger8):
    cam, r, _ = _render(menger8, "quality")
    hit = r["node"] != svo.MISS
    # LOD off: quality = scale_exp2 at the leaf's parent = 2^-depth
    assert np.array_equal(r["shade"][hit][:, 0], np.full(hit.sum(), F(2.0 ** (-8 / 4))))
    assert np.all(r["shade"][~hit][:, :3] == np.array([0.6, 0.3, 0.3], np.float32))


def _wavelength_rgb(loops):
    w = (loops.astype(np.float32) - F(0)) / (F(390) - F(0)) * (F(780) - F(380)) + F(380)
    r = np.zeros_like(w); g = np.zeros_like(w); b = np.zeros_like(w)
    m = (w >= 380) & (w <= 440); r[m] = -(w[m] - F(440)) / F(60); b[m] = 1
    m = (w > 440) & (w <= 490); g[m] = (w[m] - F(440)) / F(50); b[m] = 1
    m = (w > 490) & (w <= 510); g[m] = 1; b[m] = -(w[m] - F(510)) / F(20)
    m = (w > 510) & (w <= 580); r[m] = (w[m] - F(510)) / F(70); g[m] = 1
    m = (w > 580) & (w <= 645); r[m] = 1; g[m] = -(w[m] - F(645)) / F(65)
    m = (w > 645) & (w <= 780); r[m] = 1
    f = np.zeros_like(w)
    m = (w >= 380) & (w <= 420); f[m] = F(0.3) + F(0.7) * (w[m] - F(380)) / F(40)
    m = (w > 420) & (w <= 701); f[m] = 1
    m = (w > 701) & (w <= 780); f[m] = F(0.3) + F(0.7) * (F(780) - w[m]) / F(79)

    def adj(c):
        v = np.round(F(255) * np.power(c * f, F(0.8))).astype(np.int64)
        return np.where(c == 0, 0, np.clip(v, 0, 255)).astype(np.float32) * (F(1) / F(255))
    return np.stack([adj(r), adj(g), adj(b)], -1)


def test_iterations_colour_map(menger8):
    cam, r, _ = _render(menger8, "iterations")
    want = _wavelength_rgb(r["loops"])
    assert np.array_equal(r["shade"][..., :3], want)
    assert len(np.unique(r["loops"])) > 10


def test_phong(menger8):
    import oracle
    cam, r, inc = _render(menger8, "phong")
    hit = r["node"] != svo.MISS
    d = oracle.ray_dirs(cam)[hit]
    n = (_unpack(inc["normals"][r["node"][hit]])[:, :3] * K - F(0.5)) * F(2.0)
    n = n / np.sqrt((n[:, :1] * n[:, :1] + n[:, 1:2] * n[:, 1:2]) + n[:, 2:3] * n[:, 2:3])
    col = _unpack(inc["colors"][r["node"][hit]])[:, :3] * K
    lv = np.array([-120.0, 0.0, 0.0], np.float32) - r["center"][hit]
    lv = lv / np.sqrt((lv[:, :1] ** 2 + lv[:, 1:2] ** 2) + lv[:, 2:3] ** 2)
    ndl = np.maximum(F(0), ((n * lv)[:, 0] + (n * lv)[:, 1]) + (n * lv)[:, 2])
    rf = (F(2) * n) * ndl[:, None] - lv
    rf = rf / np.sqrt((rf[:, :1] ** 2 + rf[:, 1:2] ** 2) + rf[:, 2:3] ** 2)
    rdv = np.maximum(F(0), ((rf * -d)[:, 0] + (rf * -d)[:, 1]) + (rf * -d)[:, 2])
    spec = np.maximum(F(0), np.power(rdv, F(10))) * (ndl > 0)
    o = col * F(0.5) + (col * ndl[:, None] + (spec * F(0.2))[:, None]) * F(0.5)
    t = np.clip(o, 0, 1)
    want = t * t * (F(3) - F(2) * t)
    assert np.allclose(r["shade"][hit][:, :3], want, rtol=0, atol=2e-6)
    assert np.all(r["shade"][~hit][:, :3] == np.array([0.4, 0.6, 0.4], np.float32))


def test_svo_depth_is_depth_over_13_without_lod(menger8):
    cam, r, _ = _render(menger8, "svo_depth")
    hit = r["node"] != svo.MISS
    # LOD off: every hit is a leaf at the sponge's depth 8
    assert np.all(r["shade"][hit][:, :3] == F(8) / F(13))
    assert np.all(r["shade"][~hit][:, :3] == 0.0)


# Wavelength (nm) -> RGB of the published WavelengthToRGB algorithm (gamma
# 0.8, intensity ramps below 420 and above 701 nm), as 8-bit values, at
# wavelengths that the map reaches exactly from values v on [0, 400]
# (w = 380 + v): the band edges 380 / 580 / 780 and points inside each band.
WAVELENGTH_RGB = {
    380: (97, 0, 97),     # violet end: r = b = 1 at 30 % intensity
    430: (61, 0, 255),    # r = 1/6
    455: (0, 97, 255),    # g = 0.3
    480: (0, 213, 255),   # g = 0.8
    505: (0, 255, 84),    # b = 0.25
    530: (94, 255, 0),    # r = 2/7
    580: (255, 255, 0),   # yellow
    605: (255, 173, 0),   # g = 40/65
    680: (255, 0, 0),
    705: (248, 0, 0),     # intensity 0.3 + 0.7 * 75/79
    780: (97, 0, 0),      # red end at 30 % intensity
}


def test_colour_map_matches_published_wavelength_values():
    import oracle
    for w, rgb in WAVELENGTH_RGB.items():
        got = oracle.data_point_color(float(w - 380), 0.0, 400.0) * F(255)
        assert np.array_equal(np.round(got).astype(int), np.array(rgb)), (w, got)
    # outside [380, 780] nm the map is black
    assert not np.any(oracle.data_point_color(-10.0, 0.0, 400.0))
    assert not np.any(oracle.data_point_color(410.0, 0.0, 400.0))


def test_phong_away_from_the_light_is_the_ambient_half(menger8):
    """Where the normal faces away from the light (n.l <= 0), diffuse and
    specular terms vanish: the pixel is smoothstep(colour / 2); everywhere
    else it is no darker than that."""
    cam, r, inc = _render(menger8, "phong")
    hit = r["node"] != svo.MISS
    col = _unpack(inc["colors"][r["node"][hit]])[:, :3].astype(np.float64) / 255.0
    n = _unpack(inc["normals"][r["node"][hit]])[:, :3].astype(np.float64) / 255.0 * 2.0 - 1.0
    lv = np.array([-120.0, 0.0, 0.0]) - r["center"][hit].astype(np.float64)
    away = np.einsum("ij,ij->i", n, lv) < -1e-3 * np.linalg.norm(lv, axis=1)
    base = np.clip(col * 0.5, 0, 1)
    base = base * base * (3 - 2 * base)
    got = r["shade"][hit][:, :3].astype(np.float64)
    assert away.sum() > 50 and (~away).sum() > 50
    assert np.allclose(got[away], base[away], atol=1e-5)
    assert np.all(got >= base - 1e-5)
