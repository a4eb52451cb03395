"""Pins for the oracle's volume raymarch (shaders/raymarch.frag): texture filtering against
linear fields, the slab test against known distances, and whole frames against closed forms of
front-to-back compositing (a uniform medium gives a geometric series), exact identities (an
empty volume renders as the background; E parallel to B gives no Poynting glow) and the signed
palette's branch. The camera is the binding's orbit camera (host-only arithmetic)."""
import ctypes as C
import math

import numpy as np
import pytest

from em_lab_amd.emfdtd import Camera

N = 16          # cube lattice: world box [-1, 1]^3
W = H = 9       # odd sizes: the centre pixel's ray is the view axis


def _settings(**kw):
    s = dict(mode=1, signed_axis=2, gain_e=1.0, gain_b=1.0, density=1.0, emission=1.0, step_voxels=0.75,
             material_opacity=0.45, exposure=1.25, clip_min=(0.0, 0.0, 0.0), clip_max=(1.0, 1.0, 1.0))
    s.update(kw)
    return s


def _textures(e=(0.0, 0.0, 0.0), b=(0.0, 0.0, 0.0), sigma=0.0, eps=1.0, n=N):
    E = np.zeros((n ** 3, 4), np.float32)
    B = np.zeros((n ** 3, 4), np.float32)
    E[:, :3] = e
    E[:, 3] = sigma
    B[:, :3] = b
    B[:, 3] = eps
    return E, B


def _frame(oracle, E, B, cam=None, n=N, **kw):
    cam = cam or Camera(yaw_deg=180.0, aspect=1.0)  # eye at (-4, 0, 0), looking along +x
    args = oracle.render_args((n, n, n), _settings(**kw), cam.inverse_view_projection(), cam.position(), W, H)
    return oracle.render_volume(args, E, B), args


def _display(c, exposure, dither):
    c = np.float32(1.0) - np.exp(-np.float64(c) * exposure)
    return np.power(c, 1.0 / 2.2) + dither


def test_texture_linear_field_and_edge_clamp(oracle):
    L = oracle.lib()
    nx, ny, nz = 7, 5, 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    val = (0.5 + 0.25 * x - 0.5 * y + 0.125 * z).astype(np.float32).ravel()
    tex = np.zeros((val.size, 4), np.float32)
    tex[:, 2] = val
    rng = np.random.default_rng(3)
    for _ in range(200):
        # interior: texel centres at (i + 0.5) / n, so u n - 0.5 is the lattice coordinate
        cx, cy, cz = rng.uniform(0, nx - 1), rng.uniform(0, ny - 1), rng.uniform(0, nz - 1)
        got = L.or_texture(oracle.ptr(tex), nx, ny, nz, 2, (cx + 0.5) / nx, (cy + 0.5) / ny, (cz + 0.5) / nz)
        assert got == pytest.approx(0.5 + 0.25 * cx - 0.5 * cy + 0.125 * cz, abs=2e-6)
    # clamp to edge: below the first texel centre the value is the edge texel's, at any distance
    for u in (0.0, 0.25 / nx, -0.3):
        got = L.or_texture(oracle.ptr(tex), nx, ny, nz, 2, u, 2.5 / ny, 3.5 / nz)
        assert got == pytest.approx(0.5 - 0.5 * 2 + 0.125 * 3, abs=2e-6)
    got = L.or_texture(oracle.ptr(tex), nx, ny, nz, 2, 1.2, 2.5 / ny, 3.5 / nz)
    assert got == pytest.approx(0.5 + 0.25 * (nx - 1) - 1.0 + 0.375, abs=2e-6)
    # the other channels are not mixed in
    assert L.or_texture(oracle.ptr(tex), nx, ny, nz, 0, 0.5, 0.5, 0.5) == 0.0


def test_box_hit_distances(oracle):
    L = oracle.lib()
    f3 = lambda *v: np.array(v, np.float32)  # noqa: E731
    t0, t1 = C.c_float(), C.c_float()
    L.or_box_hit(oracle.ptr(f3(0, 0, -5)), oracle.ptr(f3(0, 0, 1)), oracle.ptr(f3(-1, -1, -1)), oracle.ptr(f3(1, 1, 1)),
                 C.byref(t0), C.byref(t1))
    assert (t0.value, t1.value) == (4.0, 6.0)
    d = np.float32(1.0 / math.sqrt(2.0))
    L.or_box_hit(oracle.ptr(f3(-3, -3, 0)), oracle.ptr(f3(d, d, 0)), oracle.ptr(f3(-1, -1, -1)), oracle.ptr(f3(1, 1, 1)),
                 C.byref(t0), C.byref(t1))
    assert t0.value == pytest.approx(2 * math.sqrt(2.0), rel=1e-6) and t1.value == pytest.approx(4 * math.sqrt(2.0), rel=1e-6)
    L.or_box_hit(oracle.ptr(f3(0, 3, -5)), oracle.ptr(f3(0, 0, 1)), oracle.ptr(f3(-1, -1, -1)), oracle.ptr(f3(1, 1, 1)),
                 C.byref(t0), C.byref(t1))
    assert t0.value > t1.value  # a miss


def test_gradient_noise_range_and_mean(oracle):
    L = oracle.lib()
    v = np.array([[L.or_gradient_noise(x + 0.5, y + 0.5) for x in range(64)] for y in range(64)])
    assert v.min() >= 0.0 and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.02


def test_empty_volume_is_background(oracle):
    # zero fields and vacuum: nothing emits or absorbs, so every pixel is the pixel of a ray that
    # misses (a clip box of zero width), bit for bit
    E, B = _textures()
    for mode in range(5):
        a, _ = _frame(oracle, E, B, mode=mode)
        miss, _ = _frame(oracle, E, B, mode=mode, clip_min=(0.5, 0.0, 0.0), clip_max=(0.5, 0.0, 0.0))
        assert np.array_equal(a, miss)
    # and the background itself: a gradient over the ray's z, tone-mapped
    cam = Camera(yaw_deg=180.0, pitch_deg=30.0, aspect=1.0)
    img, args = _frame(oracle, E, B, cam=cam, clip_min=(0.5, 0.5, 0.5), clip_max=(0.5, 0.5, 0.5))
    # centre ray looks down 30 degrees: rd.z = -sin 30
    h = -0.5 * 0.5 + 0.5
    bg = np.array([0.004, 0.006, 0.014]) * (1 - h) + np.array([0.018, 0.028, 0.060]) * h
    dither = (oracle.lib().or_gradient_noise(4.5 + 17.0, 4.5 + 17.0) - 0.5) / 255.0
    assert np.allclose(img[4, 4, :3], _display(bg, 1.25, dither), atol=2e-6)


def _composite(c, a, n):
    """n steps of front-to-back compositing of per-step colour c and opacity a; the ray stops
    after the first step whose opacity passes 0.995."""
    keep = 1.0 - a
    k = n
    for i in range(1, n + 1):
        if 1.0 - keep ** i > 0.995:
            k = i
            break
    return c * (1.0 - keep ** k) / a, 1.0 - keep ** k


def _centre_steps(oracle, args):
    # eye 4 units from the box centre on the view axis: the ray crosses 2 units
    noise = oracle.lib().or_gradient_noise(4.5, 4.5)
    ds = np.float32(0.75) * np.float32(2.0 / N)
    t = np.float32(3.0) + ds * np.float32(noise)
    return int((np.float32(5.0) - t) / ds) + 1


@pytest.mark.parametrize("m", [0.05, 0.3, 1.5])
def test_uniform_field_closed_form(oracle, m):
    # a uniform |E| = m along the view axis: each step adds colour c s_vox T and opacity
    # a = 1 - exp(-k s_vox), a geometric series in the transmittance T
    E, B = _textures(e=(0.0, m * 0.6, m * 0.8))
    img, args = _frame(oracle, E, B, mode=1)
    n = _centre_steps(oracle, args)
    i_e = m * m / (1 + m * m)
    c0, c1, c2 = np.array([0.02, 0.22, 1.0]), np.array([0.05, 0.95, 1.0]), np.array([0.8, 1.0, 1.0])
    t = min(max((i_e - 0.8) / 0.2, 0.0), 1.0)
    col = (c0 * (1 - i_e) + c1 * i_e) * (1 - t * t * (3 - 2 * t)) + c2 * (t * t * (3 - 2 * t))
    emis = col * (i_e * i_e * 2.0) * 0.045
    svox = 0.75
    a = 1.0 - math.exp(-0.022 * i_e * svox)
    acc, alpha = _composite(emis * svox, a, n)
    h = 0.5  # horizontal ray
    bg = np.array([0.004, 0.006, 0.014]) * (1 - h) + np.array([0.018, 0.028, 0.060]) * h
    dither = (oracle.lib().or_gradient_noise(21.5, 21.5) - 0.5) / 255.0
    want = _display(acc + (1 - alpha) * bg, 1.25, dither)
    assert np.allclose(img[4, 4, :3], want, atol=3e-5), (img[4, 4, :3], want)


def test_metal_block_closed_form(oracle):
    # a volume of metal (sigma far above 20): the clamped texture is uniform, its gradient 0, so
    # every step is lit facing the viewer; the ray stops once opacity passes 0.995
    E, B = _textures(sigma=100.0)
    img, args = _frame(oracle, E, B, mode=0)
    n = _centre_steps(oracle, args)
    rd = np.array([1.0, 0.0, 0.0])
    nrm, v = -rd, -rd
    key = np.array([0.45, -0.35, 0.82]) / np.linalg.norm([0.45, -0.35, 0.82])
    hv = (key + v) / np.linalg.norm(key + v)
    spec = max(nrm @ hv, 0.0) ** 48 * 0.9
    surf = np.array([0.50, 0.53, 0.58]) * (0.18 + 0.55 * max(nrm @ key, 0.0) + 0.30 * 1.0) + spec
    a = 1.0 - math.exp(-0.45 * 0.75)
    acc, alpha = _composite(surf * a, a, n)
    assert alpha > 0.995
    bg = np.array([0.011, 0.017, 0.037])
    dither = (oracle.lib().or_gradient_noise(21.5, 21.5) - 0.5) / 255.0
    assert np.allclose(img[4, 4, :3], _display(acc + (1 - alpha) * bg, 1.25, dither), atol=3e-5)


def test_poynting_of_parallel_fields_is_dark(oracle):
    E0, B0 = _textures()
    empty, _ = _frame(oracle, E0, B0, mode=3)
    E, B = _textures(e=(0.3, -0.2, 0.5), b=(0.6, -0.4, 1.0))  # B = 2 E
    par, _ = _frame(oracle, E, B, mode=3)
    assert np.array_equal(par, empty)
    E, B = _textures(e=(0.3, 0.0, 0.0), b=(0.0, 0.0, 2.0))  # |E x B| = 0.6
    perp, _ = _frame(oracle, E, B, mode=3)
    assert perp[4, 4, 0] > empty[4, 4, 0] + 0.05


def test_signed_component_palette(oracle):
    E0, B0 = _textures()
    empty, _ = _frame(oracle, E0, B0, mode=4)
    pos, _ = _frame(oracle, *_textures(e=(0.0, 0.0, 0.8)), mode=4, signed_axis=2)
    neg, _ = _frame(oracle, *_textures(e=(0.0, 0.0, -0.8)), mode=4, signed_axis=2)
    other, _ = _frame(oracle, *_textures(e=(0.8, 0.0, 0.0)), mode=4, signed_axis=2)
    assert pos[4, 4, 2] > pos[4, 4, 0]  # blue palette for E z > 0
    assert neg[4, 4, 0] > neg[4, 4, 2]  # orange palette for E z < 0
    assert np.array_equal(other, empty)  # the chosen component alone is shown
