"""Pins the oracle's restatement of the reference's path tracer (oracle/pathtrace.c) with
checks that do not retype its formulas:

* tea against the published TEA test vector (key 0, plaintext 0, 32 rounds -> 41ea3a0a),
  which fixes the round function the reference's 16-round, fixed-key variant uses;
* lcg against the closed-form jump-ahead x_n = a^n x_0 + c (a^n - 1) / (a - 1) mod 2^32;
* every BRDF's sampler against its pdf and eval: the Monte Carlo estimate
  E[eval / pdf] over MAT_sample equals the quadrature of eval over the hemisphere, and
  the fraction of valid samples equals the quadrature of pdf (a wrong normalisation,
  Jacobian or lobe exponent fails these);
* light_pdf integrates to one over the sphere of directions (it is a solid-angle density
  of light_sample_dir's points), with one and with several lights;
* the renderer's closed forms: an emitting wall filling the view is rendered as its
  emission by every integrator, a view of nothing is clear_col, frame blending is
  mix(old, new, 1/(frame+1));
* direct light sampling, material sampling and MIS agree statistically where they
  estimate the same integral."""
import numpy as np
import pytest

from gpurt import tracer as T
from gpurt import workloads as W
from oracle import oracle as O


def test_tea_known_answer():
    assert O.tea(0, 0, key=[0, 0, 0, 0], rounds=32) == 0x41EA3A0A
    # the reference's variant differs from the published cipher only by key and rounds
    assert O.tea(0, 0) != O.tea(0, 0, key=[0, 0, 0, 0], rounds=16)
    assert O.tea(5, 9) == O.tea(5, 9, key=[0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E], rounds=16)


@pytest.mark.parametrize("seed", [0, 1, 0xDEADBEEF])
def test_lcg_jump_ahead(seed):
    a, c, m = 1664525, 1013904223, 1 << 32
    n = 1000
    seq = O.randf_seq(seed, n)
    for k in (1, 2, 17, 500, 1000):
        an = pow(a, k, m)
        # c (a^k - 1) / (a - 1) = c * sum_{j<k} a^j, evaluated exactly mod 2^32
        geo = sum(pow(a, j, m) for j in range(k)) % m
        x = (an * seed + c * geo) % m
        assert seq[k - 1] == np.float32((x & 0xFFFFFF) / float(1 << 24))
    assert 0.0 <= seq.min() and seq.max() < 1.0


def _hemisphere_quadrature(brdf, albedo, rough, wo, N, nt=600, nphi=1200):
    """sum over a (theta, phi) grid about N of eval(wi) sin(theta) dtheta dphi, and of pdf."""
    T_, B_ = np.cross(N, [0.3, 0.5, 0.7]), None
    T_ = T_ / np.linalg.norm(T_)
    B_ = np.cross(N, T_)
    th = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    ph = (np.arange(nphi) + 0.5) * (2 * np.pi) / nphi
    tt, pp = np.meshgrid(th, ph, indexing="ij")
    wi = (np.sin(tt)[..., None] * (np.cos(pp)[..., None] * T_ + np.sin(pp)[..., None] * B_)
          + np.cos(tt)[..., None] * np.asarray(N))
    pdf, ev = O.mat_eval(brdf, albedo, rough, wo, N, wi.reshape(-1, 3))
    dw = (np.sin(tt) * (np.pi / 2 / nt) * (2 * np.pi / nphi)).reshape(-1)
    return (ev.astype(np.float64) * dw[:, None]).sum(0), (pdf.astype(np.float64) * dw).sum()


@pytest.mark.parametrize("brdf,rough", [(0, 0.3), (0, 1.0), (1, 0.35), (1, 0.8)])
def test_material_sampler_matches_pdf_and_eval(brdf, rough):
    albedo = np.array([0.9, 0.5, 0.2], np.float32)
    N = np.array([0.0, 0.0, 1.0], np.float32)
    wo = np.array([0.4, 0.1, -0.9], np.float32)
    wo = (wo / np.linalg.norm(wo)).astype(np.float32)  # incoming, against N
    n = 400000
    ok, wi, pdf, ev = O.mat_samples(brdf, albedo, rough, wo, N, 12345, n)
    good = ok & (pdf > 0)
    est = np.where(good[:, None], ev / np.where(pdf > 0, pdf, 1)[:, None], 0.0).astype(np.float64)
    q_eval, q_pdf = _hemisphere_quadrature(brdf, albedo, rough, wo, N)
    mc = est.mean(0)
    se = est.std(0) / np.sqrt(n)
    assert np.all(np.abs(mc - q_eval) <= 4 * se + 5e-3 * q_eval), (mc, q_eval, se)
    # the density of valid samples integrates to the fraction of valid samples
    assert abs(good.mean() - q_pdf) <= 5e-3 + 4 * np.sqrt(q_pdf * (1 - q_pdf) / n), (good.mean(), q_pdf)


def _uniform_dirs(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


@pytest.mark.parametrize("several", [False, True])
def test_light_pdf_integrates_to_one(several):
    sc = W.tracer_floor(occluder=False)
    objs = [dict(o) for o in sc["objects"]]
    if several:
        cube = W._cube()
        objs.append(W._obj(cube, W._affine((0.4, 0.3, 0.5), (10, 20, 30), (-1.0, 2.5, 0.5)), (0, 0, 0),
                           emissive=(3, 3, 3)))
    n = 1 << 20
    D = _uniform_dirs(n, 3)
    for p in ([0.3, 2.9, 0.1], [-0.5, 1.5, 1.0]):
        P = np.tile(np.asarray(p, np.float32), (n, 1))
        pdf = O.light_pdf(objs, P, D).astype(np.float64)
        est = pdf.mean() * 4 * np.pi
        se = pdf.std() * 4 * np.pi / np.sqrt(n)
        assert abs(est - 1.0) <= 4 * se + 2e-3, (p, est, se)


def _render(objs, cam, w, h, **kw):
    iv, ip = T.camera(cam["eye"], cam["target"], yfov_deg=cam.get("yfov_deg", 45.0), aspect=w / h)
    p = dict(width=w, height=h, inv_view=iv, inv_proj=ip, clear_col=(0.3, 0.3, 0.3), env_light=(0, 0, 0), frame=0,
             samples=2, max_frame=256, qmc=0, max_depth=4, use_metalness=0, integrator=0, brdf=0, use_rr=1, seed=1)
    p.update(kw)
    return O.render(objs, p), p


@pytest.mark.parametrize("integrator", [0, 1, 2])
def test_emitter_filling_the_view(integrator):
    wall = W._obj(W._quad(), W._affine((50, 1, 50), (90, 0, 0), (0, 0, -5)), (0, 0, 0), emissive=(1.5, 2.5, 3.5))
    img, _ = _render([wall], {"eye": (0, 0, 5), "target": (0, 0, 0), "yfov_deg": 30}, 16, 12,
                     integrator=integrator, samples=3)
    assert np.array_equal(img[..., :3], np.broadcast_to(np.float32([1.5, 2.5, 3.5]), (12, 16, 3)))
    assert (img[..., 3] == 1).all()


def test_nothing_in_view_and_frame_blend():
    sc = W.tracer_box(sphere_level=1)
    img, p = _render(sc["objects"], {"eye": (0, 0, 100), "target": (0, 0, 200), "yfov_deg": 10}, 8, 8,
                     clear_col=(0.1, 0.2, 0.4))
    assert np.array_equal(img[..., :3], np.broadcast_to(np.float32([0.1, 0.2, 0.4]), (8, 8, 3)))
    a, p = _render(sc["objects"], sc["camera"], 12, 10, integrator=2, seed=5)
    # frame 1 blended into a zero image is new * 1/2 (0 * 1/2 + new * 1/2)
    b, _ = _render(sc["objects"], sc["camera"], 12, 10, integrator=2, seed=6, frame=1)
    blended = O.render(sc["objects"], dict(p, seed=6, frame=1), image=a)
    w = np.float32(0.5)
    assert np.array_equal(blended[..., :3], a[..., :3] * (np.float32(1) - w) + b[..., :3])


def test_integrators_agree_on_the_oracle():
    sc = W.tracer_floor(occluder=True, roughness=0.5)
    res = {}
    for integ in (0, 1, 2):
        frames = []
        for k in range(32):  # enough frames for the standard errors
            img, _ = _render(sc["objects"], sc["camera"], 16, 16, integrator=integ, samples=64, max_depth=4,
                             seed=40 + k)
            frames.append(img[..., :3].astype(np.float64).reshape(4, 4, 4, 4, 3).mean(axis=(1, 3)))
        f = np.stack(frames)
        res[integ] = (f.mean(0), f.std(0, ddof=1) / np.sqrt(len(frames)))
    for a, b in ((0, 1), (0, 2), (1, 2)):
        z = np.abs(res[a][0] - res[b][0]) / np.sqrt(res[a][1] ** 2 + res[b][1] ** 2 + 1e-12)
        assert np.quantile(z, 0.9) < 3.0 and z.max() < 4.5, (a, b, np.sort(z.ravel())[-5:])
