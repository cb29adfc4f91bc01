"""Learned preconditioner on the GPU: the BF16-MFMA U-net kernels against the PyTorch
FP32 reference (oracle/unet_ref.py), Adam training, and flexible CG with the network
and AMG fallback inside the Newton solver.

Tolerances: the forward pass rounds weights, input and every activation to BF16 exactly
where the reference does, so the outputs differ only by FP32 accumulation order and the
rare rounding flips it causes, which later layers carry on (observed relative L2 error
4e-7 to 3e-3; the reference moves as much between FP32 and FP64 accumulation), see the
forward test.  The backward pass also feeds BF16-rounded gradient
tensors to the MFMA units, which the FP32 reference does not (observed per-layer relative
L2 error 1.7e-3 to 3.8e-3, cosine >= 0.99999; bounds 4e-2 and 0.999) -- a transposed,
flipped or misplaced tap gives errors of order 1."""
import numpy as np
import pytest
import torch

import oracle.unet_ref as U
import plasticity_amd as pa
from plasticity_amd.learned import UNet, attach, sample_model

pytestmark = pytest.mark.gpu
CH = (16, 32, 64, 64)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-30))


def _field(grid, seed):
    nx, ny, nz = grid
    r = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    smooth = np.stack([np.sin(0.4 * x + 0.3 * y) * np.cos(0.2 * z), np.cos(0.3 * x - 0.5 * z),
                       np.sin(0.25 * (x + y + z))], axis=-1).reshape(-1, 3)
    return (smooth + 0.3 * r.standard_normal(smooth.shape)).astype(np.float32)


def _sparse_integer_params(seed, per_out=4):
    """Weights with `per_out` entries of +-1 per output channel (random tap and used input
    channel), zero elsewhere: with integer inputs every product and partial sum of the
    network is an integer far below 2^24, so FP32 accumulation is exact in any order and
    each BF16 rounding of an activation acts on the same exact value on both sides."""
    rng = np.random.default_rng(seed)
    ps = []
    for k, (mode, cin, cout, *_) in enumerate(U.layer_spec(CH)):
        w = np.zeros((U.ntaps(mode), cin, cout), np.float32)
        cin_used = 3 if k == 0 else cin
        cout_used = 3 if k == len(U.layer_spec(CH)) - 1 else cout
        for co in range(cout_used):
            for _ in range(per_out):
                w[rng.integers(U.ntaps(mode)), rng.integers(cin_used), co] = rng.choice([-1.0, 1.0])
        ps.append(w)
    return ps


@pytest.mark.parametrize("grid", [(13, 11, 9), (16, 17, 18), (37, 9, 10)])
def test_forward_matches_torch_reference(need_gpu, grid):
    """Both kernels of the 3x3x3 layers -- the LDS-tiled 16x16x32 kernel (tiles of
    16 x 4 x 4 voxels, partial tiles on every grid here, several along x on the last)
    and the one-wave-per-16-voxels 16x16x16 kernel -- against the torch reference.

    Exact case: sparse +-1 weights and integer inputs make every sum exact, so the device
    must equal the reference to the last bit -- a wrong tap, channel, parity or tile
    index changes whole outputs.  Random case (He-normal weights, smooth + noisy input):
    every layer output is rounded to BF16, so FP32 accumulation-order differences flip
    the rounding of single activations, and a flip in a coarse layer moves many outputs
    a little (the reference itself, accumulated in FP32 instead of FP64, moves by up to
    3e-3 relative on such grids): relative L2 error bound 1e-2."""
    print(f"[learned] forward vs torch, grid {grid}", flush=True)
    ps_int = _sparse_integer_params(7)
    x_int = np.random.default_rng(3).integers(-8, 9, size=(int(np.prod(grid)), 3)).astype(np.float32)
    ref_int = U.forward([torch.tensor(p) for p in ps_int], x_int, grid[::-1], CH, bf16=True).numpy()
    w_int = np.concatenate([p.reshape(-1) for p in ps_int])
    ys = {}
    for kind in (1, 0):
        net = UNet(grid, CH, seed=11).set_conv_kernel(kind)
        w = net.get_params()
        x = _field(grid, 1)
        ys[kind] = net.forward(x)
        net.set_params(w_int)
        y_int = net.forward(x_int)
        net.close()
        print(f"  conv kernel {kind}: exact case max |diff| {np.abs(y_int - ref_int).max():.3g} "
              f"(outputs up to {np.abs(ref_int).max():.3g})", flush=True)
        assert np.abs(ref_int).max() > 0 and np.array_equal(y_int, ref_int)
    ps = [torch.tensor(p) for p in U.split_params(w, CH)]
    ref = U.forward(ps, x, grid[::-1], CH, bf16=True).numpy()
    for kind, y in ys.items():
        e = rel(y, ref)
        print(f"  conv kernel {kind}: random case relative L2 error {e:.2e}", flush=True)
        assert np.isfinite(y).all() and np.linalg.norm(ref) > 0
        assert e <= 1e-2
    print(f"  kernel 1 vs kernel 0: {rel(ys[1], ys[0]):.2e}", flush=True)


@pytest.mark.parametrize("kind", [1, 0])
def test_loss_gradient_matches_torch_reference(need_gpu, kind):
    grid = (12, 10, 11) if kind == 0 else (21, 10, 11)
    print(f"[learned] loss gradient vs torch, grid {grid}, conv kernel {kind}", flush=True)
    net = UNet(grid, CH, seed=5).set_conv_kernel(kind)
    w = net.get_params()
    x = _field(grid, 2)
    t = _field(grid, 3)
    loss, g = net.loss_grad(x, t)
    lref, gref = U.loss_and_grad(w, x, t, grid[::-1], CH, bf16=True)
    print(f"  loss {loss:.6f} ref {lref:.6f}", flush=True)
    assert abs(loss - lref) <= 2e-3
    off = 0
    for k, p in enumerate(U.split_params(w, CH)):
        n = p.size
        a, b = g[off:off + n], gref[off:off + n]
        off += n
        if np.linalg.norm(b) == 0:
            assert np.linalg.norm(a) == 0
            continue
        e = rel(a, b)
        cs = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        print(f"  layer {k:2d}: rel err {e:.2e} cos {cs:.5f}", flush=True)
        assert e <= 4e-2 and cs >= 0.999


def test_padding_weights_stay_zero_and_training_lowers_the_loss(need_gpu):
    grid = (9, 9, 9)
    net = UNet(grid, CH, capacity=8, seed=2)
    w0 = net.get_params()
    first = U.split_params(w0, CH)[0]
    assert np.all(first[:, 3:, :] == 0)
    # samples: target = a smoothed copy of the input
    for s in range(8):
        x = _field(grid, 10 + s)
        xs = x.reshape(9, 9, 9, 3)
        t = (xs + np.roll(xs, 1, 0) + np.roll(xs, -1, 0) + np.roll(xs, 1, 2) + np.roll(xs, -1, 2)) / 5
        net.add_pair(x, t.reshape(-1, 3))
    assert net.num_pairs == 8
    l0 = net.train(1, batch=8, lr=1e-3, seed=0)
    l1 = net.train(150, batch=4, lr=1e-3, seed=1)
    l2 = net.train(20, batch=8, lr=1e-4, seed=2)
    print(f"[learned] training loss {l0:.4f} -> {l1:.4f} -> {l2:.4f}", flush=True)
    assert l2 < 0.5 * l0
    w1 = U.split_params(net.get_params(), CH)
    assert np.all(w1[0][:, 3:, :] == 0) and np.all(w1[-1][:, :, 3:] == 0)


def test_learned_fcg_converges_with_amg_fallback(need_gpu):
    n = 12
    print(f"[learned] FCG with the network on a {n}^3 sample block", flush=True)
    o = pa.default_options(eta_min=1e-8, eta_max=1e-8)
    m, wl = sample_model(100, n, 0.1, options=o)
    net = UNet((n + 1,) * 3, CH, capacity=64, seed=3)
    # samples from the AMG solves of the first load step
    attach(m, net, "log", 1)
    for k in (1, 2):
        conv, hist, kry, res = m.newton(k / wl.nsteps, tol=wl.newton_tol, maxiter=25)
        assert conv
        m.commit()
    assert net.num_pairs > 0
    net.train(100, batch=4, lr=1e-3)
    attach(m, net, "precondition")
    conv, hist, kry, res = m.newton(3 / wl.nsteps, tol=wl.newton_tol, maxiter=25)
    used = [res.solver_used[i] for i in range(len(kry))]
    print(f"  newton {len(hist)} krylov {kry} solver {used}", flush=True)
    assert conv
    assert set(used) <= {5, 6}
    attach(m, None, "off")


def _solve_its(m, lam, rtol, mode, net=None, strict=False):
    attach(m, net, mode) if mode != "amg" else attach(m, None, "off")
    m.assemble()
    m.subtract_external(lam)
    m.impose_dirichlet(lam)
    dU, it, used, relres = m.linear_solve(rtol, strict=strict)
    attach(m, None, "off")
    return it, used, relres


def test_trained_network_gains_over_the_method_without_it(need_gpu):
    """Train on samples logged from two sample blocks, then on a held-out block count the
    GCR iterations that reduce the residual of the first Newton system 100-fold: the
    trained network's preconditioner needs fewer than the same GCR with block-Jacobi
    alone (no network) and fewer than with the untrained network."""
    n = 16
    grid = (n + 1,) * 3
    net = UNet(grid, CH, capacity=2000, seed=1)
    for seed in (2, 3):
        m, wl = sample_model(seed, n, 0.15, options=pa.default_options())
        attach(m, net, "log", 1)
        for k in (1, 2, 3):
            conv, *_ = m.newton(k / wl.nsteps, tol=wl.newton_tol, maxiter=25)
            assert conv
            m.commit()
        attach(m, None, "off")
        m.close()
    print(f"[learned] gain test: {net.num_pairs} samples", flush=True)
    loss = net.train(400, batch=4, lr=1e-3, seed=0)
    print(f"  mean training loss {loss:.3f}", flush=True)
    fresh = UNet(grid, CH, seed=1)
    m, wl = sample_model(100, n, 0.15, options=pa.default_options(cg_itmax=400))
    lam = 1 / wl.nsteps
    its = {}
    for label, mode, nn in [("trained", "precondition", net), ("untrained", "precondition", fresh),
                            ("jacobi", "jacobi", None), ("amg", "amg", None)]:
        it, used, relres = _solve_its(m, lam, 1e-2, mode, nn)
        its[label] = it
        print(f"  {label:9s}: {it:3d} its to rtol 1e-2 (solver {used}, relres {relres:.2e})", flush=True)
        if label != "amg":  # GCR stops on ||r|| <= rtol ||b||; PCG on the preconditioned norm
            assert used == 5 and relres <= 1e-2 * 1.0001
    assert its["trained"] < its["jacobi"]
    assert its["trained"] < its["untrained"]


def _plastic_system(n, precision):
    """Held-out sample block of seed 100 after two committed AMG load steps; the caller
    solves the first Newton system of load step 3 (the first plastic one)."""
    m, wl = sample_model(100, n, 0.15, options=pa.default_options(eta_min=1e-8, eta_max=1e-8,
                                                                  precision=precision))
    for k in (1, 2):
        conv, *_ = m.newton(k / wl.nsteps, tol=wl.newton_tol, maxiter=25)
        assert conv
        m.commit()
    return m, 3 / wl.nsteps


def test_config3_held_out_seed_100_on_the_128_node_grid(need_gpu):
    """The shipped network (plasticity_amd/weights, trained on samples logged from 64^3-node
    blocks of seeds 2-20) on the configuration's own 128^3-node grid (127^3 elements,
    6.2M DOF) of held-out seed 100: the first plastic Newton system solved to rtol 1e-8 by
    flexible GCR with the network (AMG takes over when GCR slows down), against FP32-IR AMG
    and the same GCR with block-Jacobi alone."""
    import os
    n = 127
    weights = os.path.join(os.path.dirname(pa.__file__), "weights", "unet_16_32_64_64.npz")
    net = UNet((n + 1,) * 3, CH).load(weights)
    out = {}
    for label, mode, nn, prec in [("learned GCR + AMG", "precondition", net, "fp64"),
                                  ("block-Jacobi GCR + AMG", "jacobi", None, "fp64"),
                                  ("FP32-IR AMG", "amg", None, "fp32_ir")]:
        m, lam = _plastic_system(n, prec)
        out[label] = _solve_its(m, lam, 1e-8, mode, nn, strict=True)  # raises unless converged
        m.close()
        it, used, relres = out[label]
        print(f"[config 3] {label:24s}: {it:4d} Krylov its (solver {used}), true relres {relres:.2e}", flush=True)
    # each solve met its own stopping test at rtol 1e-8 (GCR: ||r|| <= rtol ||b||; the AMG
    # hand-over and PCG: the preconditioned norm, so the true residual may sit a decade higher)
    assert out["learned GCR + AMG"][1] in (5, 6)
    assert all(relres <= 1e-6 for _, _, relres in out.values())


def test_error_logging_and_corrected_cycle_modes(need_gpu):
    """Attach modes 4 and 5 (FP32-IR only): error samples are kept only from solves that
    converge, at most 16 per solve; the V-cycle corrected by a network converges to the
    same Newton solution as the plain FP32-IR path (tight tolerances); both modes are
    refused under FP64."""
    n = 12
    print(f"[learned] error logging and corrected cycle on a {n}^3 sample block", flush=True)
    o64 = pa.default_options(eta_min=1e-10, eta_max=1e-10)
    m64, wl = sample_model(100, n, 0.1, options=o64)
    net = UNet((n + 1,) * 3, CH, capacity=400, seed=3)
    for mode in ("log_error", "correct"):
        with pytest.raises(Exception):
            attach(m64, net, mode)
    m64.close()
    o = pa.default_options(eta_min=1e-10, eta_max=1e-10, precision="fp32_ir")
    runs = {}
    for label in ("plain", "log", "correct"):
        m, wl = sample_model(100, n, 0.1, options=o)
        if label == "log":
            attach(m, net, "log_error", 1)
        elif label == "correct":
            net.train(50, batch=4, lr=1e-4, seed=0)
            attach(m, net, "correct")
        kry, used = [], []
        for k in (1, 2):
            conv, hist, kr, res = m.newton(k / wl.nsteps, tol=1e-10, maxiter=25)
            assert conv
            m.commit()
            kry += list(kr)
            used += [res.solver_used[i] for i in range(len(kr))]
        if label == "log":
            npairs = net.num_pairs
            print(f"  logged {npairs} error samples over {len(kry)} solves ({sum(kry)} Krylov its)", flush=True)
            assert 0 < npairs <= 16 * len(kry)
            assert npairs == sum(min(k + 1, 16) for k, u in zip(kry, used) if u == 3)
            net.set_loss(1)
        print(f"  {label}: krylov {kry} solver {used}", flush=True)
        if label == "correct":
            assert set(used) <= {10, 3, 7}
        runs[label] = m.U.copy()
        attach(m, None, "off")
        m.close()
    e = rel(runs["correct"], runs["plain"])
    print(f"  U corrected vs plain: {e:.2e}", flush=True)
    assert e <= 1e-8
