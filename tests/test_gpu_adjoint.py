"""Path-replay adjoint on the GPU against the oracle's FP64 adjoint of the
same detached-sampling estimator (itself pinned by finite differences in
test_oracle_adjoint.py).  Gradient relative error is ||g_gpu - g_fp64|| /
||g_fp64|| over all texels of a target, bound 1e-3.  FP32 atomics differ from
the FP64 reference by accumulation rounding and by the path-replay
subtraction L_rem = L - sum(terms) in float."""
import numpy as np
import pytest

from firefly_amd import abi, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(gpu, sd):
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials, textures=sd.textures)
    env = abi.EnvMap(gpu, sd.env)
    return sc, env, abi.make_camera(sd.inv_vp)


def _forward(gpu, sd, sc, env, cam, spp_batch=1):
    import torch
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    paths = torch.zeros((sd.spp, sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    abi.render_paths(gpu, sc, env, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, film, paths,
                     spp_batch=spp_batch)
    return film, paths


def _adjoint(gpu, sd, sc, env, cam, paths, d, flags=0, spp_batch=1):
    import torch
    dd = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    dd[..., :3] = torch.from_numpy(d).cuda()
    g_env = torch.zeros(sd.env.shape, dtype=torch.float32, device="cuda")
    g_tex = [torch.zeros(t.shape, dtype=torch.float32, device="cuda") for t in sd.textures]
    abi.adjoint(gpu, sc, env, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, paths, dd, env_grad=g_env,
                tex_grads=g_tex, flags=flags, spp_batch=spp_batch)
    torch.cuda.synchronize()
    return g_env.cpu().numpy(), [g.cpu().numpy() for g in g_tex]


def _loss_grad(sd):
    """d loss / d image of an L2 loss against the image of perturbed
    parameters (oracle renders)."""
    ps = O.PreparedScene(sd)
    film, _ = ps.render()
    img = film[..., :3] / film[..., 3:]
    tex, env = scenes.perturbed(sd, seed=1)
    sd2 = scenes.SceneData(**{**sd.__dict__, "textures": tex, "env": env})
    f2, _ = O.PreparedScene(sd2).render()
    target = f2[..., :3] / f2[..., 3:]
    return (2.0 * (img - target) / img.size).astype(np.float32), ps


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _textured(size=96, spp=4):
    return scenes.textured_spheres_scene(width=size, height=size, spp=spp, max_depth=4, tex_size=64, env_w=128,
                                         env_h=64, segments=16, rings=8)


def test_render_paths_matches_render(gpu):
    import torch
    sd = _textured(64, 3)
    sc, env, cam = _setup(gpu, sd)
    film, paths = _forward(gpu, sd, sc, env, cam)
    ref = torch.zeros_like(film)
    abi.render(gpu, sc, env, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, ref)
    torch.cuda.synchronize()
    assert torch.equal(film, ref)
    # the film is the in-order float sum of the stored path radiances
    acc = torch.zeros_like(film[..., :3])
    for s in range(sd.spp):
        acc = acc + paths[s, ..., :3]
    assert torch.equal(acc, film[..., :3])


@pytest.mark.parametrize("spp_batch", [1, 2])
def test_adjoint_matches_fp64_oracle(gpu, spp_batch):
    sd = _textured()
    d, ps = _loss_grad(sd)
    sc, env, cam = _setup(gpu, sd)
    _, paths = _forward(gpu, sd, sc, env, cam, spp_batch=spp_batch)
    g_env, g_tex = _adjoint(gpu, sd, sc, env, cam, paths, d, spp_batch=spp_batch)
    _, r_env, r_tex = ps.adjoint(d)
    assert _rel(g_env, r_env) <= 1e-3
    assert _rel(g_tex[0], r_tex[0]) <= 1e-3


def test_adjoint_roughness_and_albedo_textures(gpu):
    sd, _ = scenes.multiview_scene(views=2, side=64, spp=4, max_depth=5, subdiv=3, tex_size=32, env_w=128,
                                   env_h=64)
    sd.env = scenes.sky_env(128, 64)
    d, ps = _loss_grad(sd)
    sc, env, cam = _setup(gpu, sd)
    _, paths = _forward(gpu, sd, sc, env, cam)
    g_env, g_tex = _adjoint(gpu, sd, sc, env, cam, paths, d)
    _, r_env, r_tex = ps.adjoint(d)
    assert _rel(g_tex[0], r_tex[0]) <= 1e-3
    assert _rel(g_tex[1], r_tex[1]) <= 1e-3
    assert _rel(g_env, r_env) <= 1e-3


def test_deterministic_reduction(gpu):
    sd = _textured(64, 4)
    d, _ = _loss_grad(sd)
    sc, env, cam = _setup(gpu, sd)
    _, paths = _forward(gpu, sd, sc, env, cam)
    a_env, a_tex = _adjoint(gpu, sd, sc, env, cam, paths, d, flags=abi.FF_ADJOINT_DETERMINISTIC)
    b_env, b_tex = _adjoint(gpu, sd, sc, env, cam, paths, d, flags=abi.FF_ADJOINT_DETERMINISTIC)
    assert np.array_equal(a_env, b_env) and np.array_equal(a_tex[0], b_tex[0])
    f_env, f_tex = _adjoint(gpu, sd, sc, env, cam, paths, d)
    assert _rel(f_env, a_env) <= 1e-5
    assert _rel(f_tex[0], a_tex[0]) <= 1e-5


def test_env_gradient_euler_identity(gpu):
    """The estimator is homogeneous of degree 1 in the environment texels:
    sum(env * dL/denv) = sum(d * image)."""
    import torch
    sd = _textured(64, 2)
    sc, env, cam = _setup(gpu, sd)
    film, paths = _forward(gpu, sd, sc, env, cam)
    img = (film[..., :3] / film[..., 3:]).cpu().numpy().astype(np.float64)
    d = np.random.default_rng(2).normal(size=img.shape).astype(np.float32)
    g_env, _ = _adjoint(gpu, sd, sc, env, cam, paths, d)
    lhs = float((g_env.astype(np.float64) * sd.env).sum())
    rhs = float((d * img).sum())
    assert abs(lhs - rhs) <= 1e-4 * float(np.abs(d * img).sum())


def test_batched_adjoint_shares_gradients(gpu):
    """Several views of one scene in one replay accumulate into shared
    gradient buffers: equal to the sum of the single-view adjoints."""
    import torch
    sd, vps = scenes.multiview_scene(views=3, side=40, spp=2, max_depth=4, subdiv=3, tex_size=32, env_w=64,
                                     env_h=32)
    sc, env, _ = _setup(gpu, sd)
    cams = [abi.make_camera(v) for v in vps]
    rng = np.random.default_rng(4)
    ds = [rng.normal(size=(sd.height, sd.width, 3)).astype(np.float32) for _ in cams]
    films = [torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda") for _ in cams]
    paths = [torch.zeros((sd.spp, sd.height, sd.width, 4), dtype=torch.float32, device="cuda") for _ in cams]
    items = [(sc, env, c, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, f) for c, f in zip(cams, films)]
    abi.render_batch(gpu, items, path_radiance=paths, flags=abi.FF_RENDER_USE_GRAPH)
    ref_env = np.zeros(sd.env.shape, np.float64)
    ref_tex = [np.zeros(t.shape, np.float64) for t in sd.textures]
    for c, p, d in zip(cams, paths, ds):
        e, t = _adjoint(gpu, sd, sc, env, c, p, d)
        ref_env += e
        for k in range(len(t)):
            ref_tex[k] += t[k]
    g_env = torch.zeros(sd.env.shape, dtype=torch.float32, device="cuda")
    g_tex = [torch.zeros(t.shape, dtype=torch.float32, device="cuda") for t in sd.textures]
    dd = []
    for d in ds:
        x = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
        x[..., :3] = torch.from_numpy(d).cuda()
        dd.append(x)
    abi.adjoint_batch(gpu, items, paths, dd, env_grads=[g_env] * 3, tex_grads=[g_tex] * 3,
                      flags=abi.FF_RENDER_USE_GRAPH)
    torch.cuda.synchronize()
    assert _rel(g_env.cpu().numpy(), ref_env) <= 1e-5
    for k in range(len(g_tex)):
        assert _rel(g_tex[k].cpu().numpy(), ref_tex[k]) <= 1e-5


def test_autograd_wrapper(gpu):
    import torch
    from firefly_amd import torch_fn
    sd = _textured(48, 2)
    sc, env, cam = _setup(gpu, sd)
    R = torch_fn.Renderer(gpu, sc, env, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed,
                          deterministic=True)
    tex = torch.tensor(sd.textures[0], device="cuda", requires_grad=True)
    envp = torch.tensor(sd.env, device="cuda", requires_grad=True)
    target = torch.full((sd.height, sd.width, 3), 0.3, device="cuda")
    img = torch_fn.render(tex, envp, R)
    loss = ((img - target) ** 2).mean()
    loss.backward()
    d = (2.0 * (img.detach() - target) / img.numel()).cpu().numpy()
    _, paths = _forward(gpu, sd, sc, env, cam)
    g_env, g_tex = _adjoint(gpu, sd, sc, env, cam, paths, d, flags=abi.FF_ADJOINT_DETERMINISTIC)
    assert np.array_equal(tex.grad.cpu().numpy(), g_tex[0])
    assert np.array_equal(envp.grad.cpu().numpy(), g_env)
