"""The wavefront path tracer (gpurt_scene_* / gpurt_render) against the oracle's per-pixel
restatement of the reference's ray-generation shader (oracle/pathtrace.c).

Both draw the same random numbers per pixel (tea / lcg seeded by pixel and frame seed) and
cast rays with bit-identical FP32 semantics, so a pixel differs only where a last-bit
difference of cos / sin / pow (device library vs C library) flips a discrete decision of
its path (a hit near an edge, a Russian-roulette draw next to its threshold).  The tests
require nearly every pixel to agree to 1e-4 relative; the first-hit G-buffer uses no
transcendental function and must agree bit for bit.  light_pdf through the light BVH is
compared with the reference's loop over every light triangle.  A statistical test checks
that direct light sampling, material sampling and MIS converge to the same image where
they estimate the same integral."""
import os

import numpy as np
import pytest
import torch

import gpurt
from gpurt import tracer as T
from gpurt import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _params(sc, w, h, **kw):
    cam = sc["camera"]
    iv, ip = T.camera(cam["eye"], cam["target"], yfov_deg=cam["yfov_deg"], aspect=w / h)
    p = dict(width=w, height=h, inv_view=iv, inv_proj=ip, clear_col=(0.3, 0.3, 0.3), env_light=(0.0, 0.0, 0.0),
             frame=0, samples=4, max_frame=256, qmc=0, max_depth=6, use_metalness=0, integrator=0, brdf=0,
             use_rr=1, seed=11)
    p.update(kw)
    return p


def _gpu_render(scene, p, image=None, gbuffer=False):
    kw = {k: p[k] for k in ("samples", "max_depth", "integrator", "brdf", "use_rr", "clear_col", "env_light",
                            "frame", "seed", "qmc", "max_frame", "use_metalness")}
    out = scene.render(p["width"], p["height"], p["inv_view"], p["inv_proj"], image=image, gbuffer=gbuffer, **kw)
    torch.cuda.synchronize()
    return out


def _agree(img, ref, frac=0.99, rtol=1e-4):
    img, ref = np.asarray(img)[..., :3], np.asarray(ref)[..., :3]
    assert np.isfinite(img).all()
    close = np.abs(img - ref) <= rtol * (1.0 + np.abs(ref))
    per_pixel = close.all(axis=-1)
    print(f"pixels agreeing: {per_pixel.mean():.5f}; mean {img.mean():.6g} vs {ref.mean():.6g}", flush=True)
    assert per_pixel.mean() >= frac, f"only {per_pixel.mean():.4f} of pixels agree"
    # the mean image agrees closely even with the few diverged paths
    assert abs(img.mean() - ref.mean()) <= 2e-2 * abs(ref.mean()) + 1e-6


@pytest.fixture(scope="module")
def box(ctx):
    sc = W.tracer_box(sphere_level=2)
    return sc, T.Scene(ctx, sc["objects"])


@pytest.mark.parametrize("brdf", [0, 1])
@pytest.mark.parametrize("integrator", [0, 1, 2])
def test_render_matches_oracle(box, integrator, brdf):
    sc, scene = box
    p = _params(sc, 40, 32, integrator=integrator, brdf=brdf, samples=4, max_depth=6)
    print("render integrator", integrator, "brdf", brdf, flush=True)
    img, g = _gpu_render(scene, p, gbuffer=True)
    ref, rg = O.render(sc["objects"], p, gbuffer=True)
    _agree(img.cpu().numpy(), ref)
    for a, b in zip(g, rg):
        assert np.array_equal(a.cpu().numpy(), b)  # first hits use no transcendental
    assert (img[..., 3] == 1).all()


def test_render_frames_qmc_metalness(box):
    """Frame 1 (random jitter, blended into frame 0's image), Hammersley jitter, metalness,
    an environment light and no Russian roulette."""
    sc, scene = box
    p0 = _params(sc, 36, 36, integrator=2, samples=2, env_light=(0.2, 0.3, 0.4), use_metalness=1, use_rr=0,
                 max_depth=4)
    img = _gpu_render(scene, p0)
    ref = O.render(sc["objects"], p0)
    _agree(img.cpu().numpy(), ref)
    # the image is blended in place: keep the previous frame for the oracle
    p1 = dict(p0, frame=1, seed=12)
    prev = img.cpu().numpy()
    img1 = _gpu_render(scene, p1, image=img)
    assert img1.data_ptr() == img.data_ptr()
    ref1 = O.render(sc["objects"], p1, image=prev)
    _agree(img1.cpu().numpy(), ref1)
    p2 = dict(p0, frame=3, qmc=1, max_frame=8, seed=13)
    prev = img1.cpu().numpy()
    img2 = _gpu_render(scene, p2, image=img1)
    ref2 = O.render(sc["objects"], p2, image=prev)
    _agree(img2.cpu().numpy(), ref2)


def test_camera_miss_is_clear_colour(ctx, dev):
    """No geometry in view: every pixel is clear_col exactly (rt.rgen:645-650)."""
    sc = W.tracer_box(sphere_level=1)
    scene = T.Scene(ctx, sc["objects"])
    iv, ip = T.camera((0, 0, 100), (0, 0, 200), yfov_deg=10)
    img = scene.render(8, 8, iv, ip, samples=2, clear_col=(0.25, 0.5, 0.75))
    torch.cuda.synchronize()
    assert torch.equal(img[..., :3].cpu(), torch.tensor([0.25, 0.5, 0.75]).expand(8, 8, 3))


def _light_queries(sc, n, seed):
    rng = np.random.default_rng(seed)
    P = rng.uniform(-4.5, 4.5, (n, 3)).astype(np.float32)
    D = rng.normal(size=(n, 3))
    # half of the directions aim at points of the emitting objects
    lights = [o for o in sc["objects"] if np.any(np.asarray(o["emissive"]) > 0)]
    tgt = []
    for k in range(n // 2):
        o = lights[k % len(lights)]
        v = o["verts"][:, :3][o["indices"][rng.integers(len(o["indices"]))]].astype(np.float64)
        b = rng.dirichlet([1, 1, 1])
        q = b @ v
        m = np.asarray(o["model"], np.float64)
        tgt.append(m[:3, :3] @ q + m[:3, 3])
    D[: n // 2] = np.asarray(tgt) - P[: n // 2]
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    return P, D.astype(np.float32)


def test_light_pdf_matches_reference_loop(box, dev):
    sc, scene = box
    P, D = _light_queries(sc, 20000, 1)
    got = scene.light_pdf(torch.from_numpy(P).to(dev), torch.from_numpy(D).to(dev))
    torch.cuda.synchronize()
    want = O.light_pdf(sc["objects"], P, D)
    got = got.cpu().numpy()
    assert (want > 0).sum() > 5000
    # one light of two triangles: at most two terms, so the sums agree bit for bit
    assert np.array_equal(got, want)


def test_light_pdf_two_small_lights_exact(ctx, dev):
    """Two emitting objects with 14 triangles in all: light_pdf runs the reference's loop
    in its order, so the sums equal the oracle's bit for bit."""
    sc = W.tracer_box(sphere_level=1)
    objs = [dict(o) for o in sc["objects"]]
    objs[7]["emissive"] = np.array([0.5, 0.5, 0.5], np.float32)   # block (12 tris)
    scene = T.Scene(ctx, objs)
    P, D = _light_queries({"objects": objs}, 20000, 3)
    got = scene.light_pdf(torch.from_numpy(P).to(dev), torch.from_numpy(D).to(dev))
    torch.cuda.synchronize()
    want = O.light_pdf(objs, P, D)
    assert (want > 0).sum() > 5000
    assert np.array_equal(got.cpu().numpy(), want)


def test_light_pdf_several_lights(ctx, dev):
    """Three emitting objects with different triangle counts (per-light normalisation);
    94 light triangles, so the light BVH answers."""
    sc = W.tracer_box(sphere_level=1)
    objs = [dict(o) for o in sc["objects"]]
    objs[6]["emissive"] = np.array([1.0, 2.0, 3.0], np.float32)   # glossy sphere (80 tris)
    objs[7]["emissive"] = np.array([0.5, 0.5, 0.5], np.float32)   # block (12 tris)
    sc2 = {"objects": objs, "camera": sc["camera"]}
    scene = T.Scene(ctx, objs)
    P, D = _light_queries(sc2, 20000, 2)
    got = scene.light_pdf(torch.from_numpy(P).to(dev), torch.from_numpy(D).to(dev))
    torch.cuda.synchronize()
    want = O.light_pdf(objs, P, D)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=2e-6, atol=0)
    p = _params(sc2, 32, 32, integrator=2, samples=2)
    _agree(_gpu_render(scene, p).cpu().numpy(), O.render(objs, p))


@pytest.mark.parametrize("brdf,occluder", [(0, True), (1, False)])
def test_integrators_estimate_the_same_image(ctx, brdf, occluder):
    """On a scene where every path has at most one non-black bounce, direct light
    sampling, material sampling and MIS are unbiased estimators of the same image: their
    block means agree within their Monte Carlo error."""
    sc = W.tracer_floor(occluder=occluder, roughness=0.5)
    scene = T.Scene(ctx, sc["objects"])
    w = h = 32
    means = {}
    for integ in (0, 1, 2):
        frames = []
        # 32 independent frames: enough degrees of freedom for the standard errors
        for k in range(32):
            p = _params(sc, w, h, integrator=integ, brdf=brdf, samples=64, max_depth=4, use_rr=1, seed=100 + k)
            img = _gpu_render(scene, p)[..., :3].cpu().numpy().astype(np.float64)
            frames.append(img.reshape(h // 8, 8, w // 8, 8, 3).mean(axis=(1, 3)))
        frames = np.stack(frames)
        means[integ] = (frames.mean(0), frames.std(0, ddof=1) / np.sqrt(frames.shape[0]))
    for a, b in ((0, 1), (0, 2), (1, 2)):
        (ma, sa), (mb, sb) = means[a], means[b]
        z = np.abs(ma - mb) / np.sqrt(sa ** 2 + sb ** 2 + 1e-12)
        assert np.quantile(z, 0.95) < 3.0 and z.max() < 4.5, (a, b, np.sort(z.ravel())[-5:])
        assert abs(ma.mean() - mb.mean()) <= 0.02 * ma.mean()


def test_render_rejects_bad_input(ctx, dev, box):
    sc, scene = box
    iv, ip = T.camera((0, 0, 14), (0, 0, 0))
    with pytest.raises(gpurt.GpurtError, match="ReSTIR"):
        scene.render(8, 8, iv, ip, integrator=3)
    with pytest.raises(gpurt.GpurtError):
        scene.render(8, 8, iv, ip, brdf=2)
    with pytest.raises(gpurt.GpurtError):
        scene.render(8, 8, iv, ip, samples=0)
    objs = [dict(o) for o in sc["objects"]]
    bad = dict(objs[0])
    bad["indices"] = np.array([[0, 1, 99]], np.uint32)
    bad["verts"] = bad["verts"][:4]
    with pytest.raises(gpurt.GpurtError, match="index"):
        T.Scene(ctx, objs[1:] + [bad])


def _fixture_scene(name):
    """A reference scene from its committed fixtures: positions, triangles and placement
    (gltf_<name>.npz), normals, material factors and camera (gltf_<name>_shading.npz)."""
    fx = np.load(os.path.join(GOLDEN, f"gltf_{name}.npz"))
    sh = np.load(os.path.join(GOLDEN, f"gltf_{name}_shading.npz"))
    objs = []
    for k in range(fx["xf"].shape[0]):
        v = np.zeros((fx[f"obj{k}_pos"].shape[0], 12), np.float32)
        v[:, 0:3] = fx[f"obj{k}_pos"]
        v[:, 4:7] = sh[f"obj{k}_nrm"]
        m = np.eye(4)
        m[:3] = fx["xf"][k].reshape(3, 4)
        objs.append({"verts": v, "indices": fx[f"obj{k}_tris"], "model": m, "albedo": sh["albedo"][k],
                     "emissive": sh["emissive"][k], "metal_rough": sh["metal_rough"][k]})
    cam = {"eye": sh["camera_eye"], "target": sh["camera_target"], "yfov_deg": float(sh["camera_yfov_deg"])}
    return {"objects": objs, "camera": cam}


@pytest.mark.parametrize("name", ["cbox", "mis_test"])
def test_reference_scene_renders_match_oracle(ctx, name):
    """The reference's own scenes (media/cbox, media/mis_test) rendered by every integrator."""
    sc = _fixture_scene(name)
    scene = T.Scene(ctx, sc["objects"])
    for integ in (0, 1, 2):
        p = _params(sc, 48, 36, integrator=integ, samples=2, max_depth=5)
        print(name, "integrator", integ, flush=True)
        img = _gpu_render(scene, p)
        _agree(img.cpu().numpy(), O.render(sc["objects"], p))


@pytest.mark.parametrize("integrator", [0, 1, 2])
def test_render_independent_of_pixel_groups(box, integrator):
    """Pixel groups are rendered concurrently on their own streams; every pixel draws the
    same random numbers and casts the same rays whatever the grouping, so images are
    identical bit for bit (odd sizes leave a short last group)."""
    sc, scene = box
    p = _params(sc, 37, 23, integrator=integrator, samples=3, max_depth=5, seed=21)
    kw = {k: p[k] for k in ("samples", "max_depth", "integrator", "brdf", "use_rr", "clear_col", "env_light",
                            "frame", "seed", "qmc", "max_frame", "use_metalness")}
    imgs = []
    for conc in (1, 3, 8):
        img, g = scene.render(p["width"], p["height"], p["inv_view"], p["inv_proj"], gbuffer=True, concurrency=conc,
                              **kw)
        torch.cuda.synchronize()
        imgs.append((img.cpu(), [x.cpu() for x in g]))
    for img, g in imgs[1:]:
        assert torch.equal(img, imgs[0][0])
        assert all(torch.equal(x, y) for x, y in zip(g, imgs[0][1]))
