"""Reverse-time gradients (vs_adjoint_forward / vs_adjoint_backward) on the GPU.

Bars:
* FP64 adjoint against central finite differences of the FP64 forward step
  (h = 1e-4 relative on the float32 inputs, integer geometry checked to be
  unchanged by the perturbation): relative error <= 1e-4 per component
  (components below 1e-3 of the largest of their kind: absolute 1e-4 of that
  largest), for area, depth scale and wall admittance.  The same derivative
  is also taken by finite differences of the CPU oracle's forward step.
* FP32 adjoint against the FP64 adjoint: relative L2 <= 2e-3 per kind.
* The differentiable forward's probe matches the ordinary step: FP64 within
  1e-12 relative L2, FP32 within 1e-5.
* Checkpoint spacing and batching do not change the result (bit-identical).
* The FP32 reverse pass runs on chip by default (scaled adjoints Kp = kP M,
  Lx = gX lambda_x, Ly = gY lambda_y in the layout of the FP32 step kernel):
  its forward probe is bit-identical to the FP32 step kernel's; its gradient
  matches the FP32 per-step kernels (relative L2 <= 1e-4 per kind) and the
  FP64 adjoint (<= 2e-3); batching is bit-identical; checkpoint spacing and
  shorter record blocks change only the grouping of double sums (<= 1e-12).
"""
import numpy as np
import pytest

import oracle as orc
import vocalsim
from vocalsim import Solver, VocalSimError, _abi, tubes

pytestmark = pytest.mark.gpu

FD_REL = 1e-4
H_REL = 1e-4


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def base_tube(seed=3, num_steps=1200):
    sp = tubes.vowel_tube(seed, num_steps=num_steps, source=tubes.SOURCE_GAUSSIAN)
    n = sp.num_sections
    rng = np.random.default_rng(seed)
    return sp.replace(depth_scale=rng.uniform(0.85, 1.15, n).astype(np.float32),
                      mu3d=np.full(n, 0.005, dtype=np.float32))


def loss_weights(n, seed=0):
    rng = np.random.default_rng(100 + seed)
    return rng.standard_normal(n) * np.hanning(n)


def forward_J(spec, g, precision=64):
    with Solver(spec, precision=precision) as s:
        s.run()
        return float(np.dot(g, s.probe())), s.field("cell_class")


def oracle_J(spec, g):
    p, _ = orc.OracleTube(spec, 64).run()
    return float(np.dot(g, p))


def perturbed(spec, kind, sec, sign):
    field = {"area": "areas_cm2", "scale": "depth_scale", "mu": "mu3d"}[kind]
    arr = np.array(getattr(spec, field), dtype=np.float32)
    v0 = arr[sec]
    arr[sec] = np.float32(float(v0) * (1.0 + sign * H_REL))
    return spec.replace(**{field: arr}), float(arr[sec])


def adjoint_grads(specs, gs, precision=64, checkpoint_every=512, adjoint_kernel=0, block=0):
    with Solver(specs, precision=precision) as s:
        s.set_option(_abi.OPT_ADJOINT_KERNEL, adjoint_kernel)
        s.set_option(_abi.OPT_ADJOINT_BLOCK, block)
        probes = s.adjoint_forward(checkpoint_every)
        ga, gsc, gm = s.adjoint_backward(gs)
    return probes, ga, gsc, gm


SECTIONS = [4, 19, 33]


def test_adjoint_fp64_matches_central_fd(hip_device):
    spec = base_tube()
    n = spec.num_steps
    g = loss_weights(n)
    print(f"[adjoint fd] vowel3 gauss steps={n} sections={SECTIONS}", flush=True)
    _, ga, gsc, gm = adjoint_grads([spec], [g])
    grads = {"area": ga[0], "scale": gsc[0], "mu": gm[0]}
    _, cls0 = forward_J(spec, g)
    for kind in ("area", "scale", "mu"):
        big = np.abs(grads[kind]).max()
        assert big > 0, kind
        for sec in SECTIONS:
            sp_p, vp = perturbed(spec, kind, sec, +1)
            sp_m, vm = perturbed(spec, kind, sec, -1)
            jp, clp = forward_J(sp_p, g)
            jm, clm = forward_J(sp_m, g)
            assert np.array_equal(clp, cls0) and np.array_equal(clm, cls0), "perturbation moved a wall"
            fd = (jp - jm) / (vp - vm)
            adj = grads[kind][sec]
            err = abs(adj - fd)
            print(f"  {kind}[{sec}] adjoint={adj:.10e} fd={fd:.10e} rel={err / max(abs(fd), 1e-300):.2e}")
            if abs(fd) >= 1e-3 * big:
                assert err <= FD_REL * abs(fd), (kind, sec, adj, fd)
            else:
                assert err <= FD_REL * big, (kind, sec, adj, fd)


def test_adjoint_matches_fd_of_the_oracle(hip_device):
    spec = base_tube(seed=5, num_steps=900)
    g = loss_weights(spec.num_steps, seed=1)
    _, ga, gsc, gm = adjoint_grads([spec], [g])
    for kind, grad in (("area", ga[0]), ("scale", gsc[0]), ("mu", gm[0])):
        sec = 12
        sp_p, vp = perturbed(spec, kind, sec, +1)
        sp_m, vm = perturbed(spec, kind, sec, -1)
        fd = (oracle_J(sp_p, g) - oracle_J(sp_m, g)) / (vp - vm)
        print(f"[adjoint vs oracle fd] {kind}[{sec}] adjoint={grad[sec]:.10e} fd={fd:.10e}", flush=True)
        assert abs(grad[sec] - fd) <= FD_REL * max(abs(fd), 1e-3 * np.abs(grad).max())


def test_forward_probe_matches_ordinary_step(hip_device):
    spec = base_tube(num_steps=700)
    for precision, tol in ((64, 1e-12), (32, 1e-5)):
        with Solver(spec, precision=precision) as s:
            p_adj = s.adjoint_forward(128)[0]
            s.run()
            p = s.probe()
        r = rel_l2(p_adj, p)
        print(f"[adjoint forward] precision={precision} rel_l2={r:.3e}", flush=True)
        assert r <= tol


def test_fp32_adjoint_matches_fp64_adjoint(hip_device):
    spec = base_tube(num_steps=1000)
    g = loss_weights(spec.num_steps)
    _, a64, s64, m64 = adjoint_grads([spec], [g], 64)
    _, a32, s32, m32 = adjoint_grads([spec], [g], 32)
    for kind, x32, x64 in (("area", a32, a64), ("scale", s32, s64), ("mu", m32, m64)):
        r = rel_l2(x32[0], x64[0])
        print(f"[fp32 vs fp64 adjoint] {kind} rel_l2={r:.3e}", flush=True)
        assert r <= 2e-3, kind


def test_checkpoint_spacing_and_batching_are_invisible(hip_device):
    a = base_tube(seed=3, num_steps=800)
    b = base_tube(seed=4, num_steps=650).replace(lengths_cm=np.full(44, 15.0 / 44, dtype=np.float32))
    ga_ = loss_weights(a.num_steps, 0)
    gb_ = loss_weights(b.num_steps, 1)
    pa, A1, S1, M1 = adjoint_grads([a], [ga_], checkpoint_every=512)
    _, A2, S2, M2 = adjoint_grads([a], [ga_], checkpoint_every=97)
    for x, y in ((A1, A2), (S1, S2), (M1, M2)):
        assert np.array_equal(x[0], y[0])
    pb, B1, SB1, MB1 = adjoint_grads([b], [gb_], checkpoint_every=200)
    pab, AB, SAB, MAB = adjoint_grads([a, b], [ga_, gb_], checkpoint_every=200)
    assert np.array_equal(pab[0], pa[0]) and np.array_equal(pab[1], pb[0])
    for x, y in ((AB[0], A1[0]), (SAB[0], S1[0]), (MAB[0], M1[0]), (AB[1], B1[0]), (SAB[1], SB1[0]),
                 (MAB[1], MB1[0])):
        assert np.array_equal(x, y)


def test_two_mass_source_rejected(hip_device):
    spec = tubes.vowel_tube(0, num_steps=100)
    with Solver(spec, precision=64) as s:
        with pytest.raises(VocalSimError, match="table source"):
            s.adjoint_forward(64)
        with pytest.raises(VocalSimError, match="vs_adjoint_forward"):
            s.adjoint_backward([np.zeros(100)])


def test_torch_function_backward(hip_device):
    import torch
    from vocalsim.adjoint import tube_probe

    spec = base_tube(num_steps=900)
    target_spec = spec.replace(areas_cm2=spec.areas_cm2 * np.float32(1.02))
    with Solver(target_spec, precision=64) as s:
        s.run()
        target = torch.tensor(s.probe(), device="cuda")
    dev = "cuda"
    area = torch.tensor(spec.areas_cm2, dtype=torch.float64, device=dev)[None].requires_grad_(True)
    scale = torch.tensor(spec.depth_scale, dtype=torch.float64, device=dev)[None].requires_grad_(True)
    mu = torch.tensor(spec.mu3d, dtype=torch.float64, device=dev)[None].requires_grad_(True)
    p = tube_probe(area, scale, mu, [spec], 64, 256)
    loss = 0.5 * ((p[0] - target) ** 2).sum()
    loss.backward()
    resid = (p[0] - target).detach().cpu().numpy()
    _, ga, gsc, gm = adjoint_grads([spec], [resid], checkpoint_every=256)
    assert np.array_equal(area.grad[0].cpu().numpy(), ga[0])
    assert np.array_equal(scale.grad[0].cpu().numpy(), gsc[0])
    assert np.array_equal(mu.grad[0].cpu().numpy(), gm[0])
    # finite difference of the quadratic loss in one area
    sec = 20
    sp_p, vp = perturbed(spec, "area", sec, +1)
    sp_m, vm = perturbed(spec, "area", sec, -1)
    t = target.cpu().numpy()

    def L(sp):
        with Solver(sp, precision=64) as s:
            s.run()
            return 0.5 * float(((s.probe() - t) ** 2).sum())

    fd = (L(sp_p) - L(sp_m)) / (vp - vm)
    g = float(area.grad[0, sec])
    print(f"[torch adjoint] dL/dA[{sec}] adjoint={g:.10e} fd={fd:.10e}", flush=True)
    assert abs(g - fd) <= FD_REL * abs(fd)


@pytest.mark.parametrize("precision", [32, 64])
def test_torch_function_device_and_host_tensors_agree(hip_device, precision):
    """GPU tensors (the library writes the probe traces into device memory and
    reads the gradient traces from it) and CPU tensors (host copies) give the
    same probe and gradients bit for bit, and the probe is the ABI's."""
    import torch
    from vocalsim.adjoint import tube_probe

    specs = [base_tube(seed=s, num_steps=700) for s in (1, 2)]
    res = {}
    for dev in ("cpu", "cuda"):
        print(f"[torch adjoint] tensors on {dev}, precision {precision}", flush=True)
        T = lambda name: torch.tensor(np.stack([getattr(sp, name) for sp in specs]), dtype=torch.float64,
                                      device=dev).requires_grad_(True)
        A, D, M = T("areas_cm2"), T("depth_scale"), T("mu3d")
        p = tube_probe(A, D, M, specs, precision, 256)
        assert p.device.type == dev and p.dtype == torch.float64 and p.shape == (2, 700)
        w = torch.tensor(np.stack([loss_weights(700, k) for k in range(2)]), dtype=torch.float64, device=dev)
        (p * w).sum().backward()
        res[dev] = [x.detach().cpu().numpy() for x in (p, A.grad, D.grad, M.grad)]
    for name, a, b in zip(("probe", "area", "depth scale", "mu3d"), res["cpu"], res["cuda"]):
        assert np.array_equal(a, b), name
    with Solver(specs, precision=precision, threads=1024) as s:
        probes = s.adjoint_forward(256)
    assert np.array_equal(res["cuda"][0], np.stack(probes))


def test_log_spectrum_loss_gradient_four_tubes(hip_device):
    """The gradient workload: log-magnitude spectra (0-5 kHz, 1024 bins, FP64
    DFT of the probe) of 4 tubes against targets from 4 other seeds, FP64
    adjoint, 512-step checkpoints; central FD on one parameter of each kind."""
    import torch
    from vocalsim.adjoint import dft_basis, log_magnitude, tube_probe

    N = 2000
    fs = 44100.0 * tubes.SRATE_MUL

    def spec_of(seed):
        sp = tubes.vowel_tube(seed, num_steps=N, source=tubes.SOURCE_GAUSSIAN)
        return sp.replace(depth_scale=np.ones(sp.num_sections, np.float32),
                          mu3d=np.full(sp.num_sections, 0.005, np.float32))

    specs = [spec_of(k) for k in range(4)]
    basis = dft_basis(N, fs, np.linspace(0.0, 5000.0, 1024))
    with Solver([spec_of(64 + k) for k in range(4)], precision=64) as s:
        s.run()
        target = log_magnitude(torch.tensor(np.stack([s.probe(i) for i in range(4)]), device="cuda"), basis)

    def loss_of(probe):
        return ((log_magnitude(probe, basis) - target) ** 2).mean()

    dev = "cuda"
    A = torch.tensor(np.stack([sp.areas_cm2 for sp in specs]), dtype=torch.float64, device=dev).requires_grad_(True)
    D = torch.tensor(np.stack([sp.depth_scale for sp in specs]), dtype=torch.float64, device=dev).requires_grad_(True)
    M = torch.tensor(np.stack([sp.mu3d for sp in specs]), dtype=torch.float64, device=dev).requires_grad_(True)
    loss = loss_of(tube_probe(A, D, M, specs, 64, 512))
    loss.backward()

    def L(perturbed_specs):
        with Solver(perturbed_specs, precision=64) as s:
            s.run()
            p = torch.tensor(np.stack([s.probe(i) for i in range(4)]), device="cuda")
        return float(loss_of(p))

    for kind, grad, tube, sec in (("area", A.grad, 1, 10), ("scale", D.grad, 2, 30), ("mu", M.grad, 3, 5)):
        sp_p, vp = perturbed(specs[tube], kind, sec, +1)
        sp_m, vm = perturbed(specs[tube], kind, sec, -1)
        fd = (L(specs[:tube] + [sp_p] + specs[tube + 1:]) - L(specs[:tube] + [sp_m] + specs[tube + 1:])) / (vp - vm)
        g = float(grad[tube, sec])
        print(f"[log-spectrum loss] {kind}[{tube},{sec}] adjoint={g:.10e} fd={fd:.10e}", flush=True)
        assert abs(g - fd) <= FD_REL * abs(fd), (kind, g, fd)


def test_onchip_forward_probe_is_the_fp32_step_kernel(hip_device):
    a = base_tube(seed=3, num_steps=700)
    b = base_tube(seed=6, num_steps=650)
    with Solver([a, b], precision=32) as s:
        p_adj = s.adjoint_forward(200)
    with Solver([a, b], precision=32, threads=1024, kernel=_abi.K_FAST) as s:
        s.run()
        p_run = [s.probe(0), s.probe(1)]
    print(f"[on-chip forward] max |diff| = {max(np.abs(p_adj[i] - p_run[i]).max() for i in range(2)):.3e}", flush=True)
    assert np.array_equal(p_adj[0], p_run[0]) and np.array_equal(p_adj[1], p_run[1])


def test_onchip_reverse_matches_per_step_kernels(hip_device):
    spec = base_tube(seed=7, num_steps=1100)
    g = loss_weights(spec.num_steps, seed=2)
    p1, a1, s1, m1 = adjoint_grads([spec], [g], 32, 256, adjoint_kernel=0)
    p2, a2, s2, m2 = adjoint_grads([spec], [g], 32, 256, adjoint_kernel=1)
    print(f"[on-chip vs per-step fp32] probe rel_l2={rel_l2(p1[0], p2[0]):.3e}", flush=True)
    assert rel_l2(p1[0], p2[0]) <= 1e-5
    for kind, x, y in (("area", a1, a2), ("scale", s1, s2), ("mu", m1, m2)):
        r = rel_l2(x[0], y[0])
        print(f"[on-chip vs per-step fp32] {kind} rel_l2={r:.3e}", flush=True)
        assert np.all(np.isfinite(x[0])) and r <= 1e-4, kind


def test_onchip_reverse_blocks_spacing_and_batching(hip_device):
    a = base_tube(seed=3, num_steps=800)
    b = base_tube(seed=4, num_steps=650).replace(lengths_cm=np.full(44, 15.0 / 44, dtype=np.float32))
    ga_ = loss_weights(a.num_steps, 0)
    gb_ = loss_weights(b.num_steps, 1)
    ref = adjoint_grads([a], [ga_], 32, 512)
    for label, kw in (("spacing 97", dict(checkpoint_every=97)), ("block 60", dict(checkpoint_every=512, block=60)),
                      ("spacing 300 block 128", dict(checkpoint_every=300, block=128))):
        got = adjoint_grads([a], [ga_], 32, **kw)
        assert np.array_equal(got[0][0], ref[0][0]), label
        for x, y in zip(got[1:], ref[1:]):
            r = rel_l2(x[0], y[0])
            print(f"[on-chip {label}] rel_l2={r:.3e}", flush=True)
            assert r <= 1e-12, label
    pb, B1, SB1, MB1 = adjoint_grads([b], [gb_], 32, 200)
    pa, A1, S1, M1 = adjoint_grads([a], [ga_], 32, 200)
    pab, AB, SAB, MAB = adjoint_grads([a, b], [ga_, gb_], 32, 200)
    assert np.array_equal(pab[0], pa[0]) and np.array_equal(pab[1], pb[0])
    for x, y in ((AB[0], A1[0]), (SAB[0], S1[0]), (MAB[0], M1[0]), (AB[1], B1[0]), (SAB[1], SB1[0]),
                 (MAB[1], MB1[0])):
        assert np.array_equal(x, y)


def test_onchip_slabs_kept_across_handles(hip_device):
    """After vs_destroy the on-chip pass's large slabs stay with the library
    and the next handle takes them: its probes and gradients equal those of
    freshly allocated slabs bit for bit, also for a smaller batch that takes
    larger kept slabs; vs_release_cached_memory returns them to the device."""
    import torch
    vocalsim.release_cached_memory()
    specs = [base_tube(seed=s, num_steps=700) for s in (3, 4, 5, 6)]
    gs = [loss_weights(sp.num_steps, seed=i) for i, sp in enumerate(specs)]
    fresh = adjoint_grads(specs, gs, 32, 512)              # allocates, then leaves its slabs kept
    again = adjoint_grads(specs, gs, 32, 512)              # takes the kept slabs
    small_kept = adjoint_grads(specs[:2], gs[:2], 32, 512)
    torch.cuda.synchronize()
    free_kept = torch.cuda.mem_get_info()[0]
    vocalsim.release_cached_memory()
    free_released = torch.cuda.mem_get_info()[0]
    print(f"[kept slabs] released {(free_released - free_kept) / 2**20:.0f} MiB", flush=True)
    assert free_released - free_kept >= 64 << 20
    small_fresh = adjoint_grads(specs[:2], gs[:2], 32, 512)
    vocalsim.release_cached_memory()
    for got, ref, n in ((again, fresh, 4), (small_kept, small_fresh, 2)):
        for x, y in zip(got, ref):
            for b in range(n):
                assert np.array_equal(x[b], y[b])
