"""Learned preconditioner: host side of the U-net (pfem_net_* in include/pfem.h).

The network is a bias-free 3D convolutional U-net over the structured node grid of a
box mesh (4 levels, channels 16/32/64/64, BF16 MFMA kernels with FP32 accumulation).
It is trained on (residual, V-cycle correction) samples that the solver logs from its
own AMG-preconditioned solves, with Adam and the scale-invariant loss
1 - cos(N(r), z), and then preconditions flexible CG with an AMG fallback on stall.

    net = UNet((n + 1,) * 3, capacity=2000)
    collect_samples(net, seeds=range(2, 21), n=63)      # Newton solves, samples logged
    net.train(steps=4000, batch=4, lr=1e-3)
    model.attach_net(net, mode="precondition")          # FCG + network, AMG fallback

The network and its training are additions with no counterpart in the reference
solver (lagrangian/src/Solver.jl:235-290 uses AMG-preconditioned CG only).
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _lib as L

CHANNELS = (16, 32, 64, 64)
MODES = {"off": 0, "log": 1, "precondition": 2, "jacobi": 3, "log_error": 4, "correct": 5}
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_16_32_64_64.npz")


def _f32(a, n):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class UNet:
    """Device-resident network for a node grid (nx, ny, nz), x fastest."""

    def __init__(self, grid, channels=CHANNELS, capacity=0, seed=0):
        self.grid = tuple(int(g) for g in grid)
        self.channels = tuple(int(c) for c in channels)
        self.nvox = int(np.prod(self.grid))
        d = L.NetDesc(self.grid[0], self.grid[1], self.grid[2], (C.c_int32 * 4)(*self.channels),
                      int(capacity), int(seed))
        h = C.c_void_p()
        L.check(L.lib().pfem_net_create(C.byref(d), C.byref(h)), "pfem_net_create")
        self._h = h
        self.nparams = int(L.lib().pfem_net_num_params(h))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().pfem_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_loss(self, kind):
        """Training loss: 0 = 1 - cos(N(x), t), 1 = ||N(x) - t||^2 / ||t||^2."""
        L.check(L.lib().pfem_net_set_loss(self._h, int(kind)), "net_set_loss")
        return self

    def set_conv_kernel(self, kind):
        """3x3x3 layers on the LDS-tiled 16x16x32 MFMA kernel (1, default) or the
        one-wave-per-16-voxels 16x16x16 kernel (0)."""
        L.check(L.lib().pfem_net_set_conv_kernel(self._h, int(kind)), "net_set_conv_kernel")
        return self

    def get_params(self):
        w = np.empty(self.nparams, dtype=np.float32)
        L.check(L.lib().pfem_net_get_params(self._h, L.ptr(w)), "net_get_params")
        return w

    def set_params(self, w):
        w = _f32(w, self.nparams)
        L.check(L.lib().pfem_net_set_params(self._h, L.ptr(w)), "net_set_params")

    def forward(self, x):
        x = _f32(x, 3 * self.nvox)
        out = np.empty(3 * self.nvox, dtype=np.float32)
        L.check(L.lib().pfem_net_forward(self._h, L.ptr(x), L.ptr(out)), "net_forward")
        return out.reshape(-1, 3)

    def loss_grad(self, x, target):
        x = _f32(x, 3 * self.nvox)
        t = _f32(target, 3 * self.nvox)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        L.check(L.lib().pfem_net_loss_grad(self._h, L.ptr(x), L.ptr(t), C.byref(loss), L.ptr(g)), "net_loss_grad")
        return loss.value, g

    def add_pair(self, r, z):
        L.check(L.lib().pfem_net_add_pair(self._h, L.ptr(_f32(r, 3 * self.nvox)), L.ptr(_f32(z, 3 * self.nvox))),
                "net_add_pair")

    @property
    def num_pairs(self):
        n = C.c_int64()
        L.check(L.lib().pfem_net_num_pairs(self._h, C.byref(n)), "net_num_pairs")
        return n.value

    def train(self, steps, batch=4, lr=1e-3, seed=0):
        loss = C.c_double()
        L.check(L.lib().pfem_net_train(self._h, int(steps), int(batch), float(lr), int(seed), C.byref(loss)),
                "net_train")
        return loss.value

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, w=self.get_params(), channels=np.asarray(self.channels))

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            if tuple(int(c) for c in f["channels"]) != self.channels:
                raise ValueError("channel configuration of the weight file differs")
            self.set_params(f["w"])
        return self


def attach(model, net, mode="precondition", stride=1):
    """Attach `net` to a Model: 'log' samples from AMG solves, 'precondition' uses the
    network (coarse correction + block-Jacobi) in flexible CG with AMG fallback,
    'jacobi' the same flexible CG without the network, 'off' detaches."""
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    L.check(L.lib().pfem_attach_net(model._h, None if net is None else net._h, m, int(stride)), "attach_net")
    model._net = net if m else None


def grid_of(mesh):
    """(nx+1, ny+1, nz+1) node counts of a box mesh (node = x + nx*(y + ny*z))."""
    if mesh.dims is None:
        raise ValueError("the learned preconditioner needs a structured box mesh")
    nx, ny, nz = mesh.dims[:3]
    return (nx + 1, ny + 1, nz + 1)


def sample_model(seed, n, inclusion, options=None):
    """A workload-1-style Model for sample collection / evaluation."""
    from .model import J2Material, Model
    from .workloads import workload
    wl = workload(3, n=n, seed=seed, inclusion=inclusion)
    mats = [J2Material(**m) for m in wl.materials]
    m = Model(wl.mesh, mats if len(mats) > 1 else mats[0], options=options, material_id=wl.material_id)
    m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    m.reset()
    m._freeze()
    return m, wl


def inclusion_fraction(seed):
    """Training meshes span inclusion fractions 0-30% (a fixed function of the seed)."""
    return 0.3 * ((seed * 7) % 11) / 10.0


def collect_samples(net, seeds, n, load_steps=(1, 2, 3), stride=2, log=print):
    """Run Newton load steps on the sample meshes with the AMG solver, logging every
    `stride`-th (residual, V-cycle correction) pair into `net`'s store."""
    from .model import ElementInversionError, default_options
    t0 = time.perf_counter()
    for s in seeds:
        m, wl = sample_model(s, n, inclusion_fraction(s), options=default_options())
        attach(m, net, "log", stride)
        for k in load_steps:
            try:
                conv, hist, kry, _ = m.newton(k / wl.nsteps, tol=wl.newton_tol, maxiter=25)
            except ElementInversionError as e:  # samples logged so far are kept
                log(f"[learned] seed {s} load step {k}: {e}")
                break
            if not conv:
                log(f"[learned] seed {s} load step {k}: Newton did not converge")
                break
            m.commit()
        attach(m, None, "off")
        log(f"[learned] seed {s} inclusions {inclusion_fraction(s):.2f}: samples {net.num_pairs} "
            f"({time.perf_counter() - t0:.1f}s)")
        m.close()
    return net.num_pairs


def collect_error_samples(net, seeds, n, load_steps=(1, 2, 3), stride=1, log=print):
    """FP32-IR Newton load steps (linear rtol 1e-8) on the sample meshes; every
    `stride`-th FP32 V-cycle application of their flexible PCG solves logs an error
    sample (attach mode 'log_error': input V(r_k), target x* - x_k - V(r_k), the error
    the cycle left, both divided by max |V(r_k)|)."""
    from .model import ElementInversionError, default_options
    t0 = time.perf_counter()
    for s in seeds:
        m, wl = sample_model(s, n, inclusion_fraction(s),
                             options=default_options(precision="fp32_ir", eta_min=1e-8, eta_max=1e-8))
        attach(m, net, "log_error", stride)
        for k in load_steps:
            try:
                conv, hist, kry, _ = m.newton(k / wl.nsteps, tol=wl.newton_tol, maxiter=25)
            except ElementInversionError as e:
                log(f"[learned] seed {s} load step {k}: {e}")
                break
            if not conv:
                log(f"[learned] seed {s} load step {k}: Newton did not converge")
                break
            m.commit()
        attach(m, None, "off")
        log(f"[learned] seed {s} inclusions {inclusion_fraction(s):.2f}: error samples {net.num_pairs} "
            f"({time.perf_counter() - t0:.1f}s)")
        m.close()
    return net.num_pairs


def evaluate_correction(net_weights, seeds, n, load_steps=(1, 2, 3), log=print):
    """Newton load steps on held-out sample meshes under FP32-IR at linear rtol 1e-8, with
    the plain FP32 V-cycle and with the V-cycle corrected by the network (attach mode
    'correct').  Returns per-seed Krylov iterations, solve seconds and fallbacks."""
    from .model import default_options
    out = []
    for s in seeds:
        row = {"seed": s, "inclusion": inclusion_fraction(s)}
        for label in ("fp32_ir", "correct"):
            m, wl = sample_model(s, n, inclusion_fraction(s),
                                 options=default_options(precision="fp32_ir", eta_min=1e-8, eta_max=1e-8))
            net = None
            if label == "correct":
                net = UNet(grid_of(wl.mesh), CHANNELS)
                net.set_params(net_weights)
                attach(m, net, "correct")
            kry = newton = 0
            t_solve = 0.0
            used = []
            for k in load_steps:
                conv, hist, kr, res = m.newton(k / wl.nsteps, tol=wl.newton_tol, maxiter=25)
                if not conv:
                    raise RuntimeError(f"seed {s} load step {k} ({label}) did not converge")
                m.commit()
                kry += sum(kr)
                newton += len(hist)
                t_solve += res.t_solve
                used += [res.solver_used[i] for i in range(len(kr))]
            row[label] = {"krylov": kry, "newton": newton, "t_solve": t_solve, "solves": len(used),
                          "plain_finish": sum(1 for u in used if u != 10) if label == "correct" else 0}
            attach(m, None, "off")
            if net is not None:
                net.close()
            m.close()
        log(f"[learned] correction eval seed {s}: FP32-IR {row['fp32_ir']['krylov']} its "
            f"{row['fp32_ir']['t_solve']:.3f}s | corrected {row['correct']['krylov']} its "
            f"{row['correct']['t_solve']:.3f}s (plain cycle finished {row['correct']['plain_finish']} of "
            f"{row['correct']['solves']} solves)")
        out.append(row)
    return out


def evaluate(net_weights, seeds, n, load_steps=(1, 2, 3), precision="fp64", log=print):
    """Newton load steps on held-out sample meshes at linear rtol 1e-8, once with the AMG
    solver and once with the learned preconditioner in flexible GCR (AMG fallback), and
    once with the same GCR without the network (block-Jacobi).  Returns per-seed totals."""
    from .model import default_options
    out = []
    for s in seeds:
        row = {"seed": s, "inclusion": inclusion_fraction(s)}
        for label in ("amg", "learned", "jacobi"):
            m, wl = sample_model(s, n, inclusion_fraction(s),
                                 options=default_options(precision=precision, eta_min=1e-8, eta_max=1e-8))
            net = None
            if label == "learned":
                net = UNet(grid_of(wl.mesh), CHANNELS)
                net.set_params(net_weights)
                attach(m, net, "precondition")
            elif label == "jacobi":
                attach(m, None, "jacobi")
            kry = newton = 0
            t_solve = 0.0
            used = []
            for k in load_steps:
                conv, hist, kr, res = m.newton(k / wl.nsteps, tol=wl.newton_tol, maxiter=25)
                if not conv:
                    raise RuntimeError(f"seed {s} load step {k} ({label}) did not converge")
                m.commit()
                kry += sum(kr)
                newton += len(hist)
                t_solve += res.t_solve
                used += [res.solver_used[i] for i in range(len(kr))]
            row[label] = {"krylov": kry, "newton": newton, "t_solve": t_solve,
                          "fallbacks": sum(1 for u in used if u == 6), "solves": len(used)}
            attach(m, None, "off")
            if net is not None:
                net.close()
            m.close()
        log(f"[learned] eval seed {s}: AMG {row['amg']['krylov']} its {row['amg']['t_solve']:.3f}s | "
            f"learned {row['learned']['krylov']} its {row['learned']['t_solve']:.3f}s, "
            f"fallbacks {row['learned']['fallbacks']}/{row['learned']['solves']} | "
            f"jacobi {row['jacobi']['krylov']} its {row['jacobi']['t_solve']:.3f}s, "
            f"fallbacks {row['jacobi']['fallbacks']}/{row['jacobi']['solves']}")
        out.append(row)
    return out


def layer_flops(grid, channels=CHANNELS):
    """MFMA operations (2 per multiply-add, padded channel counts) of each layer of one
    forward pass on a node grid: (name, taps, cin, cout, output voxels, flops).  Level
    l + 1 has ceil(n / 2) nodes per direction; the layer table is unet.cpp's."""
    c0, c1, c2, c3 = channels
    lv = [tuple(int(g) for g in grid)]
    for _ in range(3):
        lv.append(tuple((n + 1) // 2 for n in lv[-1]))
    vox = [int(np.prod(g)) for g in lv]
    table = [("in", 27, 16, c0, 0), ("e0", 27, c0, c0, 0), ("d1", 8, c0, c1, 1), ("e1", 27, c1, c1, 1),
             ("d2", 8, c1, c2, 2), ("e2", 27, c2, c2, 2), ("d3", 8, c2, c3, 3), ("e3", 27, c3, c3, 3),
             ("u2", 27, c3, c2, 2), ("u1", 27, c2, c1, 1), ("u0", 27, c1, c0, 0), ("out", 1, c0, 16, 0)]
    return [(n, t, ci, co, vox[l], 2 * vox[l] * t * ci * co) for n, t, ci, co, l in table]


def bench_net(n=100, reps=20, train_steps=10, conv_kernel=1, log=print):
    """Times the forward pass (host copies of r and N(r) included) and Adam steps on a
    grid of (n + 1)^3 nodes with random weights and random logged pairs; the kernels'
    own times come from a kernel trace of this command.  conv_kernel: 1 LDS-tiled
    16x16x32 kernel for the 3x3x3 layers, 0 the one-wave-per-16-voxels kernel."""
    grid = (n + 1,) * 3
    net = UNet(grid, CHANNELS, capacity=8, seed=0).set_conv_kernel(conv_kernel)
    rng = np.random.default_rng(0)
    x = rng.normal(size=3 * net.nvox).astype(np.float32)
    for _ in range(4):
        net.add_pair(rng.normal(size=3 * net.nvox), rng.normal(size=3 * net.nvox))
    net.forward(x)
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(x)
    t_fwd = (time.perf_counter() - t0) / reps
    net.train(1, batch=2, lr=1e-4, seed=0)
    t0 = time.perf_counter()
    net.train(train_steps, batch=2, lr=1e-4, seed=1)
    t_step = (time.perf_counter() - t0) / train_steps
    fl = sum(r[-1] for r in layer_flops(grid))
    log(f"[net] {n + 1}^3 nodes: forward {t_fwd * 1e3:.3f} ms ({fl / 1e9:.1f} GFLOP, {fl / t_fwd / 1e12:.1f} TFLOP/s "
        f"with host copies); Adam step (batch 2) {t_step * 1e3:.1f} ms")
    net.close()
    return {"forward_ms": t_fwd * 1e3, "forward_gflop": fl / 1e9, "adam_step_ms": t_step * 1e3}


def _seeds(spec):
    a, b = spec.split("-") if "-" in spec else (spec, spec)
    return list(range(int(a), int(b) + 1))


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train / evaluate the learned preconditioner")
    ap.add_argument("cmd", choices=["train", "eval", "bench", "correct"])
    ap.add_argument("--n", type=int, default=63, help="elements per side of the sample blocks")
    ap.add_argument("--seeds", default="2-20")
    ap.add_argument("--load-steps", default="1-5")
    ap.add_argument("--stride", type=int, default=1)
    ap.add_argument("--capacity", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--out", default=None)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_ir"])
    ap.add_argument("--eval-seeds", default="100-100", help="correct: held-out seeds")
    ap.add_argument("--conv-kernel", type=int, default=1, choices=[0, 1],
                    help="bench: 3x3x3 layers on the LDS-tiled 16x16x32 kernel (1) or the 16x16x16 one (0)")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.cmd == "train":
        net = UNet((a.n + 1,) * 3, CHANNELS, capacity=a.capacity, seed=0)
        collect_samples(net, _seeds(a.seeds), a.n, load_steps=_seeds(a.load_steps), stride=a.stride, log=log)
        t0 = time.perf_counter()
        chunk = max(1, a.steps // 20)
        done = 0
        while done < a.steps:
            k = min(chunk, a.steps - done)
            lr = a.lr if done < 0.8 * a.steps else 0.1 * a.lr
            loss = net.train(k, batch=a.batch, lr=lr, seed=done)
            done += k
            log(f"[learned] adam step {done}/{a.steps}: mean loss {loss:.4f} ({time.perf_counter() - t0:.1f}s)")
        net.save(a.out or a.weights)
        log(f"[learned] saved {a.out or a.weights}")
    elif a.cmd == "correct":
        # error samples from FP32-IR solves, relative-MSE training, held-out evaluation
        net = UNet((a.n + 1,) * 3, CHANNELS, capacity=a.capacity, seed=0).set_loss(1)
        collect_error_samples(net, _seeds(a.seeds), a.n, load_steps=_seeds(a.load_steps), stride=a.stride, log=log)
        t0 = time.perf_counter()
        chunk = max(1, a.steps // 10)
        done = 0
        while done < a.steps:
            k = min(chunk, a.steps - done)
            lr = a.lr if done < 0.8 * a.steps else 0.1 * a.lr
            loss = net.train(k, batch=a.batch, lr=lr, seed=done)
            done += k
            log(f"[learned] adam step {done}/{a.steps}: mean relative squared error {loss:.4f} "
                f"({time.perf_counter() - t0:.1f}s)")
        w = net.get_params()
        if a.out:
            net.save(a.out)
        net.close()
        res = evaluate_correction(w, _seeds(a.eval_seeds), a.n, load_steps=_seeds(a.load_steps), log=log)
        print(json.dumps(res))
    elif a.cmd == "bench":
        print(json.dumps(bench_net(a.n, conv_kernel=a.conv_kernel, log=log)))
    else:
        with np.load(a.weights, allow_pickle=False) as f:
            w = f["w"].copy()
        rows = evaluate(w, _seeds(a.seeds), a.n, load_steps=_seeds(a.load_steps), precision=a.precision, log=log)
        print(json.dumps(rows))


if __name__ == "__main__":
    main()
