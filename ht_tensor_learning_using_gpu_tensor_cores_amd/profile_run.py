"""Profiling driver: a fixed number of decompositions of one named workload, optionally
followed by one calibration copy of the tensor, so that rocprofv3 can attribute kernel
time (--kernel-trace --stats) and HBM counters (--pmc FETCH_SIZE / WRITE_SIZE) to the
engine's kernels. The copy kernel (copy_f32_kernel) moves exactly 8 |X| / 4 bytes, which
calibrates the counters.

  python3 -m ht_tensor_learning_using_gpu_tensor_cores_amd.profile_run --workload c2 --reps 3 --copy

Workloads (generated on the device):
  c1  3rd order 128^3, Tucker core 8^3 + 1e-3 noise, seed 0, tree {1},{2,3}, ranks 8, plain FP32
  c2  4th order 128^4, Tucker core 16^4 + 1e-4 noise, seed 1, tree {1,2},{3,4}, ranks 16,
      leaves-to-root BF16 (the benchmark workload)
  c3  3rd order 1024^3, core 32^3 with mode spectra k^-1.5 + 1e-5 noise, seed 2, tree
      {1},{2,3}, ranks 32, randomised (TSQR) BF16
  c4  4096 3rd-order tensors, mode sizes U{16..64}, Tucker ranks U{4..8}, 1e-3 noise, seed 3,
      tree {1},{2,3}: the batched entry point (one persistent launch); reports aggregate
      decompositions/s and the p50/p99 relative errors
  c5  learned per-mode sketch on 128^4 Gaussian-bump tensors, tree {1,2},{3,4}: trains the
      sketch (width --width, default 20) on --count tensors of seed 4 (generated on the
      device one at a time), evaluates the sketch loss on 200 held-out tensors of seed 5
      against the untrained Gaussian sketch, then decomposes --reps held-out tensors at
      ranks min(12, width) with the randomised method (BF16, no power iteration) using the
      learned sketch (fallback above --tol), the Gaussian test matrix, and the plain method
      (the exact comparator); --save writes the trained sketch (BF16) to the given .npz
"""
import argparse
import time


def workload(name):
    from . import tree as T
    if name == "c1":
        return dict(dims=[128] * 3, gen=("tucker", [8] * 3, 0, 1e-3), tree=T.first_vs_rest_3d(), ranks=[1] + [8] * 4,
                     method="plain", precision="fp32")
    if name == "c2":
        return dict(dims=[128] * 4, gen=("tucker", [16] * 4, 1, 1e-4), tree=T.balanced_4d(), ranks=[1] + [16] * 6,
                    method="ltr", precision="bf16")
    if name == "c3":
        return dict(dims=[1024] * 3, gen=("cp_decay", 32, 2, 1e-5), tree=T.first_vs_rest_3d(), ranks=[1] + [32] * 4,
                    method="rsvd", precision="bf16")
    raise ValueError(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--count", type=int, default=None, help="c4: number of tensors (4096); c5: training tensors (2000)")
    ap.add_argument("--width", type=int, default=20, help="c5: sketch width")
    ap.add_argument("--epochs", type=int, default=1, help="c5: training epochs")
    ap.add_argument("--save", default=None, help="c5: .npz for the trained sketch")
    ap.add_argument("--tol", type=float, default=1e-3, help="c5: residual tolerance of the Gaussian fallback")
    ap.add_argument("--method", default=None)
    ap.add_argument("--precision", default=None)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy", action="store_true", help="one calibration copy of the tensor after the runs")
    args = ap.parse_args()
    import torch
    from . import engine as E
    from . import _lib
    if args.workload == "c4":
        args.count = args.count or 4096
        return batched(args, torch, E)
    if args.workload == "c5":
        args.count = args.count or 2000
        return learned(args, torch, E)
    w = workload(args.workload)
    method = args.method or w["method"]
    prec = args.precision or w["precision"]
    h = E.default_handle(0)
    kind = w["gen"][0]
    if kind == "tucker":
        _, core, seed, noise = w["gen"]
        x = E.generate_tucker(w["dims"], core, seed=seed, noise_rel=noise, handle=h)
    else:
        _, core, seed, noise = w["gen"]
        x = E.generate_cp_decay(w["dims"], core, seed=seed, noise_rel=noise, handle=h)
    torch.cuda.synchronize()
    for i in range(args.reps + 1):
        t0 = time.perf_counter()
        r = E.decompose(x, w["dims"], w["tree"], w["ranks"], method=method, precision=prec, handle=h)
        torch.cuda.synchronize()
        print(f"{args.workload} {method}/{prec} rep {i}: {(time.perf_counter() - t0) * 1e3:.3f} ms, "
              f"rel_err {r.rel_err:.10f}", flush=True)
    if args.copy:
        lib = _lib.load()
        y = torch.empty_like(x)
        st = lib.ht_debug_copy(h.ptr, x.data_ptr(), y.data_ptr(), x.numel())
        torch.cuda.synchronize()
        print(f"calibration copy of {x.numel() * 4} bytes: status {st}", flush=True)


def batched(args, torch, E):
    import numpy as np
    dims, core = E.batch_shapes(args.count, 3)
    t0 = time.perf_counter()
    batch = E.generate_batch(dims, core, 3, 1e-3)
    torch.cuda.synchronize()
    gb = sum(int(np.prod(d)) for d in dims) * 4 / 1e9
    print(f"c4: generated {args.count} tensors ({gb:.3f} GB) in {time.perf_counter() - t0:.1f} s", flush=True)
    ranks = E.batch_ranks(core)
    for i in range(args.reps + 1):
        t0 = time.perf_counter()
        _, err, st = E.decompose_batched(batch, ranks)
        wall = time.perf_counter() - t0
        print(f"c4 rep {i}: device {st['total_ms']:.3f} ms ({args.count / st['total_ms'] * 1e3:.0f} decomps/s, "
              f"{gb / st['total_ms'] * 1e3:.1f} GB/s), wall {wall * 1e3:.1f} ms, rel_err p50 {np.percentile(err, 50):.3e} "
              f"p99 {np.percentile(err, 99):.3e} max {err.max():.3e}", flush=True)


def learned(args, torch, E):
    import numpy as np
    from . import sketch as SK
    from . import tree as T
    n, w = 128, args.width
    tr = T.balanced_4d()
    init = SK.gaussian([n] * 4, w, seed=7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sk, curve = SK.train_bumps(n, 4, tr, w, count=args.count, seed=4, epochs=args.epochs, batch=8, lr=0.05,
                               init=init)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"c5 w={w}: trained on {args.count} tensors x {args.epochs} epochs in {dt:.1f} s "
          f"({dt / (args.count * args.epochs) * 1e3:.2f} ms per tensor step); epoch losses {curve}", flush=True)
    skb = sk.bf16()
    if args.save:
        skb.save(args.save)
    t0 = time.perf_counter()
    lg = SK.eval_bumps(init, n, 4, tr, 200, seed=5)
    ll = SK.eval_bumps(skb, n, 4, tr, 200, seed=5)
    print(f"c5 w={w}: held-out sketch loss (200 tensors): gaussian mean {lg.mean():.4e} p50 {np.median(lg):.4e}; "
          f"learned(bf16) mean {ll.mean():.4e} p50 {np.median(ll):.4e}; ratio {ll.mean() / lg.mean():.3f} "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    k = min(12, w)
    ranks = [1] + [k] * 6
    for i in range(args.reps):
        x = E.generate_bumps(n, 4, 5, i)
        torch.cuda.synchronize()
        out = {}
        for name, kw in (("learned", dict(method="rsvd", precision="bf16", power_iters=0, sketch=skb,
                                          sketch_tol=args.tol)),
                         ("gaussian", dict(method="rsvd", precision="bf16", power_iters=0, oversample=w - k)),
                         ("plain", dict(method="plain", precision="fp32"))):
            try:
                r = E.decompose(x, [n] * 4, tr, ranks, **kw)
                out[name] = f"{r.rel_err:.6e} ({r.stats['total_ms']:.2f} ms"
                if name == "learned":
                    out[name] += (f", resid est {r.stats['sketch_resid']:.2e}, "
                                  f"fallbacks {r.stats['sketch_fallbacks']}")
                out[name] += ")"
            except E._lib.HTError as ex:
                out[name] = f"failed: {ex}"
        print(f"c5 tensor {i} ranks {k}: rel_err learned {out['learned']}, gaussian {out['gaussian']}, "
              f"plain {out['plain']}", flush=True)
        del x


if __name__ == "__main__":
    main()
