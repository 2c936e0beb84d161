"""GPU adjoint of the dense method against the FP64 oracle adjoint and
central finite differences of the GPU forward.

The oracle adjoint (orc_adjoint.c) is itself pinned by FP64 central
differences (test_oracle_adjoint.py). Here the GPU gradient of
L = sum(out * g) over two views, with respect to level-0 albedo and opacity,
must match it to 1e-3 relative L2 in every state mode (samples stored by
the forward on the tape, kept per lane by a second march, or fetched again
in the reverse sweep), and match FP32 central differences (h = 1e-3) of the
GPU forward on sampled voxels.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

RES = 16
W, H = 40, 30
VIEWS = 2


def _case():
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import scenes

    rng = np.random.default_rng(42)
    scene = scenes.room(801, scale=1.0, n_boxes=16, n_spheres=2)
    pos, nrm, mat = scene.world_triangles()
    views = []
    for v in range(VIEWS):
        vp, vp_inv, eye = scenes.camera(810 + v, W, H, scale=1.0)
        g = O.raycast_gbuffer(scene, pos, nrm, mat, vp, vp_inv, W, H)
        gd = {k: torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else
                                  (a.view(np.int16) if a.dtype == np.uint16 else a)).cuda() for k, a in g.items()}
        views.append(dict(g=g, gd=gd, gb=V.gbuffer_struct(gd), cam=V.camera_ubo(vp, vp_inv, eye), vp_inv=vp_inv,
                          gout=rng.normal(size=(H, W, 4))))
    albedo = rng.uniform(0.1, 0.9, (RES, RES, RES, 3)).astype(np.float32)
    opacity = rng.uniform(0.05, 0.6, (RES, RES, RES)).astype(np.float32)
    return views, albedo, opacity


TD = (0.872665, 1.2, 1.0, 1.0, 2.0, 0.5, 8.0, 0)


def _tape_layout(res=RES, extent=1.0, td=TD, width=W, height=H, cones=16):
    """The tape's documented layout: (cap, count offset, bytes). cap is the
    march's step count (float32 arithmetic, as the host counts it) plus 2."""
    f = np.float32
    vs = f(extent) / f(res)
    coef = f(2.0) * f(np.tan(f(td[0]) * f(0.5)))
    sf = f(max(td[2], 0.2))
    step, diam, n = f(0.0), vs, 0
    while step < f(td[1]):
        n += 1
        step = f(step + f(max(diam, vs)) * sf)
        diam = f(step * coef)
    cap = n + 2
    up = lambda v: (v + 255) // 256 * 256  # noqa: E731
    slots = ((width + 15) // 16) * ((height + 15) // 16) * 4 * cones
    counts = slots * cap * 64 * 16
    headers = up(counts + slots * 64)
    starts = up(headers + slots * 72)
    ctr = up(starts + width * height * 16)
    return cap, counts, ctr + 16


class GpuGrid:
    def __init__(self):
        import vk_voxel_cone_tracing_amd as V

        A = V._abi
        self.A = A
        self.h = C.c_void_p()
        A.call("vct_dense_create", C.byref(self.h), C.byref(A.DenseDesc(A.f3(0.0, 0.0, 0.0), 1.0, RES, 2, 1, 0)))
        n = C.c_size_t()
        A.call("vct_dense_texels", self.h, C.byref(n))
        self.texels = n.value

    def close(self):
        self.A.call("vct_dense_destroy", self.h)

    def forward(self, views, albedo, opacity):
        import torch

        A = self.A
        ta = torch.from_numpy(np.ascontiguousarray(albedo)).cuda()
        to = torch.from_numpy(np.ascontiguousarray(opacity)).cuda()
        A.call("vct_dense_set_level0", self.h, None, C.c_void_p(ta.data_ptr()), C.c_void_p(to.data_ptr()))
        A.call("vct_dense_mipmap", self.h, None)
        outs = []
        for v in views:
            out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
            A.call("vct_dense_trace", self.h, None, C.byref(v["gb"]), C.byref(v["cam"]),
                   C.byref(A.DenseTraceDesc(*TD)), C.c_void_p(out.data_ptr()), None)
            outs.append(out)
        torch.cuda.synchronize()
        return [o.cpu().numpy().astype(np.float64) for o in outs]

    def forward_tape(self, views, albedo, opacity):
        """vct_dense_trace_tape per view: (outputs, tapes)."""
        import torch

        A = self.A
        ta = torch.from_numpy(np.ascontiguousarray(albedo)).cuda()
        to = torch.from_numpy(np.ascontiguousarray(opacity)).cuda()
        A.call("vct_dense_set_level0", self.h, None, C.c_void_p(ta.data_ptr()), C.c_void_p(to.data_ptr()))
        A.call("vct_dense_mipmap", self.h, None)
        outs, tapes = [], []
        for v in views:
            nb = C.c_size_t()
            A.call("vct_dense_tape_bytes", self.h, C.byref(v["gb"]), C.byref(v["cam"]),
                   C.byref(A.DenseTraceDesc(*TD)), C.byref(nb))
            tape = torch.full((nb.value,), 0xA5, dtype=torch.uint8, device="cuda")  # no zero left from earlier use
            out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
            A.call("vct_dense_trace_tape", self.h, None, C.byref(v["gb"]), C.byref(v["cam"]),
                   C.byref(A.DenseTraceDesc(*TD)), C.c_void_p(out.data_ptr()), None, C.c_void_p(tape.data_ptr()),
                   nb.value)
            outs.append(out)
            tapes.append(tape)
        torch.cuda.synchronize()
        return [o.cpu().numpy().astype(np.float64) for o in outs], tapes

    def backward(self, views, flags, tapes=None, together=False):
        import torch

        A = self.A
        grad = torch.zeros(self.texels * 4, dtype=torch.float32, device="cuda")
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        gos = [torch.from_numpy(v["gout"].astype(np.float32)).cuda() for v in views]
        if together:  # vct_dense_trace_backward_tapes over all the views
            n = len(views)
            gbs = (A.GBuffer * n)(*[v["gb"] for v in views])
            cams = (A.CameraUBO * n)(*[v["cam"] for v in views])
            A.call("vct_dense_trace_backward_tapes", self.h, None, n, gbs, cams, C.byref(A.DenseTraceDesc(*TD)),
                   (C.c_void_p * n)(*[g.data_ptr() for g in gos]), (C.c_void_p * n)(*[t.data_ptr() for t in tapes]),
                   min(t.numel() for t in tapes), C.c_void_p(grad.data_ptr()), C.c_void_p(status.data_ptr()))
        for k, v in enumerate([] if together else views):
            go = gos[k]
            if tapes is not None:
                A.call("vct_dense_trace_backward_tape", self.h, None, C.byref(v["gb"]), C.byref(v["cam"]),
                       C.byref(A.DenseTraceDesc(*TD)), C.c_void_p(go.data_ptr()), C.c_void_p(tapes[k].data_ptr()),
                       tapes[k].numel(), C.c_void_p(grad.data_ptr()), C.c_void_p(status.data_ptr()))
            else:
                A.call("vct_dense_trace_backward", self.h, None, C.byref(v["gb"]), C.byref(v["cam"]),
                       C.byref(A.DenseTraceDesc(*TD)), C.c_void_p(go.data_ptr()), C.c_void_p(grad.data_ptr()), flags,
                       C.c_void_p(status.data_ptr()))
        A.call("vct_dense_mipmap_backward", self.h, None, C.c_void_p(grad.data_ptr()))
        ga = torch.zeros((RES, RES, RES, 3), dtype=torch.float32, device="cuda")
        gop = torch.zeros((RES, RES, RES), dtype=torch.float32, device="cuda")
        A.call("vct_dense_level0_grad", self.h, None, C.c_void_p(grad.data_ptr()), C.c_void_p(ga.data_ptr()),
               C.c_void_p(gop.data_ptr()))
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        return ga.cpu().numpy().astype(np.float64), gop.cpu().numpy().astype(np.float64)


def _oracle(views, albedo, opacity):
    A = O.Adjoint(RES)
    rgba = np.concatenate([albedo, opacity[..., None]], axis=-1).astype(np.float64)
    A.set_level0(rgba)
    outs, grad = [], np.zeros((RES, RES, RES, 4))
    for v in views:
        p = O.dense_params(v["vp_inv"], aperture=TD[0], max_dist=TD[1], decay=TD[4], ao_factor=TD[5],
                           diffuse_intensity=TD[6])
        outs.append(A.render(v["g"], p))
        grad += A.backward(v["g"], p, v["gout"])
    return outs, grad[..., :3], grad[..., 3]


@pytest.mark.parametrize("flags", [0, 1])
def test_gradient_matches_fp64_oracle(flags):
    views, albedo, opacity = _case()
    G = GpuGrid()
    try:
        outs = G.forward(views, albedo, opacity)
        ga, gop = G.backward(views, flags)
    finally:
        G.close()
    ref_outs, ra, rop = _oracle(views, albedo, opacity)
    for o, r in zip(outs, ref_outs):
        assert np.linalg.norm(o - r) / np.linalg.norm(r) < 1e-4
    for name, g, r in (("albedo", ga, ra), ("opacity", gop, rop)):
        err = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"flags={flags} {name}: rel L2 vs FP64 oracle {err:.2e}, nonzero {np.count_nonzero(r)}")
        assert np.count_nonzero(r) > 500
        assert err <= 1e-3, (name, err)


def test_tape_forward_equals_trace_and_tape_gradient_matches():
    """The tape forward's images equal vct_dense_trace's bit for bit; the
    tape's step counts sum to the march's sample count; the backward from the
    tape matches the FP64 oracle to 1e-3 and the kept-samples backward to
    1e-5 (same sample gradients, float atomics in another order)."""
    import torch

    views, albedo, opacity = _case()
    G = GpuGrid()
    try:
        outs = G.forward(views, albedo, opacity)
        keep_a, keep_o = G.backward(views, 0)
        touts, tapes = G.forward_tape(views, albedo, opacity)
        for o, t in zip(outs, touts):
            np.testing.assert_array_equal(o, t)
        # step counts: one byte per (wave slot, cone, lane) after the samples
        A = G.A
        for v, tape in zip(views, tapes):
            samples = torch.zeros(1, dtype=torch.int64, device="cuda")
            out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
            A.call("vct_dense_trace", G.h, None, C.byref(v["gb"]), C.byref(v["cam"]), C.byref(A.DenseTraceDesc(*TD)),
                   C.c_void_p(out.data_ptr()), C.c_void_p(samples.data_ptr()))
            tiles = ((W + 15) // 16) * ((H + 15) // 16)
            cap, off, nbytes = _tape_layout()
            assert tape.numel() == nbytes
            counts = tape[off:off + tiles * 4 * 16 * 64].cpu().numpy()
            assert counts.max() <= cap
            assert int(counts.astype(np.int64).sum()) == int(samples.item()) > 0
        ga, gop = G.backward(views, 0, tapes)
        _, tapes2 = G.forward_tape(views, albedo, opacity)
        ga2, gop2 = G.backward(views, 0, tapes2, together=True)
    finally:
        G.close()
    for g, g2 in ((ga, ga2), (gop, gop2)):  # one call over the views = one call per view
        assert np.linalg.norm(g2 - g) / np.linalg.norm(g) < 1e-5
    ref_outs, ra, rop = _oracle(views, albedo, opacity)
    for name, g, r, k in (("albedo", ga, ra, keep_a), ("opacity", gop, rop, keep_o)):
        err = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"tape {name}: rel L2 vs FP64 oracle {err:.2e}, vs kept samples "
              f"{np.linalg.norm(g - k) / np.linalg.norm(k):.2e}")
        assert err <= 1e-3, (name, err)
        assert np.linalg.norm(g - k) / np.linalg.norm(k) < 1e-5


def test_tape_refusals():
    """A tape smaller than vct_dense_tape_bytes, a misaligned one and an RGBA8
    grid are refused before any launch."""
    import torch
    import vk_voxel_cone_tracing_amd as V

    A = V._abi
    lib = A.lib()
    views, albedo, opacity = _case()
    v = views[0]
    G = GpuGrid()
    h8 = C.c_void_p()
    try:
        nb = C.c_size_t()
        A.call("vct_dense_tape_bytes", G.h, C.byref(v["gb"]), C.byref(v["cam"]), C.byref(A.DenseTraceDesc(*TD)),
               C.byref(nb))
        assert nb.value == _tape_layout()[2]
        tape = torch.zeros(nb.value + 16, dtype=torch.uint8, device="cuda")
        out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        go = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        grad = torch.zeros(G.texels * 4, dtype=torch.float32, device="cuda")
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        td = A.DenseTraceDesc(*TD)
        small = nb.value - 1
        assert lib.vct_dense_trace_tape(G.h, None, C.byref(v["gb"]), C.byref(v["cam"]), C.byref(td),
                                        C.c_void_p(out.data_ptr()), None, C.c_void_p(tape.data_ptr()), small) == 1
        assert lib.vct_dense_trace_tape(G.h, None, C.byref(v["gb"]), C.byref(v["cam"]), C.byref(td),
                                        C.c_void_p(out.data_ptr()), None, C.c_void_p(tape.data_ptr() + 4),
                                        nb.value) == 1
        assert lib.vct_dense_trace_backward_tape(G.h, None, C.byref(v["gb"]), C.byref(v["cam"]), C.byref(td),
                                                 C.c_void_p(go.data_ptr()), C.c_void_p(tape.data_ptr()), small,
                                                 C.c_void_p(grad.data_ptr()), C.c_void_p(status.data_ptr())) == 1
        A.call("vct_dense_create", C.byref(h8), C.byref(A.DenseDesc(A.f3(0.0, 0.0, 0.0), 1.0, RES, 0, 1, 0)))
        assert lib.vct_dense_trace_tape(h8, None, C.byref(v["gb"]), C.byref(v["cam"]), C.byref(td),
                                        C.c_void_p(out.data_ptr()), None, C.c_void_p(tape.data_ptr()),
                                        nb.value) == 5
        torch.cuda.synchronize()
        assert float(out.abs().sum()) == 0.0 and int(tape.sum()) == 0
    finally:
        if h8.value:
            A.call("vct_dense_destroy", h8)
        G.close()


def test_gradient_matches_central_differences():
    views, albedo, opacity = _case()
    G = GpuGrid()
    try:
        G.forward(views, albedo, opacity)
        ga, gop = G.backward(views, 0)

        def loss(a, o):
            return sum(float((out * v["gout"]).sum()) for out, v in zip(G.forward(views, a, o), views))

        rng = np.random.default_rng(7)
        h = 1e-3
        scale = max(np.abs(ga).max(), np.abs(gop).max())
        checked = 0
        for which, g in (("albedo", ga), ("opacity", gop)):
            flat = np.abs(g).reshape(-1)
            picks = np.concatenate([np.argsort(flat)[-12:], rng.choice(flat.size, 8, replace=False)])
            for e in picks:
                a, o = albedo.copy(), opacity.copy()
                tgt = a if which == "albedo" else o
                tgt.reshape(-1)[e] += h
                lp = loss(a, o)
                tgt.reshape(-1)[e] -= 2 * h
                lm = loss(a, o)
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g.reshape(-1)[e]) <= 2e-3 * scale + 2e-2 * abs(fd), (which, e, fd, g.reshape(-1)[e])
                checked += 1
        assert checked == 40
    finally:
        G.close()


def test_autograd_function_matches_c_abi_gradient():
    import torch
    from vk_voxel_cone_tracing_amd import adjoint, scenes

    views, albedo, opacity = _case()
    G = GpuGrid()
    try:
        ref_outs = G.forward(views, albedo, opacity)
        ra, rop = G.backward(views, 0)
    finally:
        G.close()
    cams = adjoint.cameras_tensor([scenes.camera(810 + v, W, H, scale=1.0) for v in range(VIEWS)])
    stack = lambda k: torch.stack([v["gd"][k] for v in views]).contiguous()  # noqa: E731
    ta = torch.from_numpy(albedo).cuda().requires_grad_(True)
    to = torch.from_numpy(opacity).cuda().requires_grad_(True)
    p = adjoint.TraceParams(*TD)
    gout = torch.from_numpy(np.stack([v["gout"] for v in views]).astype(np.float32)).cuda()
    for mode in ("tape", "keep", "recompute", "retained"):
        ta.grad = None
        to.grad = None
        out = adjoint.render(ta, to, stack("diffuse"), stack("normal"), stack("depth"), cams, (0.0, 0.0, 0.0), 1.0,
                             p, recompute=mode == "recompute", tape=mode != "keep")
        for k in range(VIEWS):
            np.testing.assert_allclose(out[k].detach().cpu().numpy(), ref_outs[k], rtol=0, atol=1e-6)
        if mode == "retained":
            # the first backward consumes the tape; the second marches again
            out.backward(gout, retain_graph=True)
            first = (ta.grad.clone(), to.grad.clone())
            ta.grad = None
            to.grad = None
            out.backward(gout)
            for g, f in zip((ta.grad, to.grad), first):
                assert float((g - f).norm() / f.norm()) < 1e-5
        else:
            out.backward(gout)
        for g, r in ((ta.grad, ra), (to.grad, rop)):
            g = g.cpu().numpy().astype(np.float64)
            assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-5, mode
