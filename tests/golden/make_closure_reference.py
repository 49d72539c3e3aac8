"""Writes the filtered-DNS reference of the closure configuration as a fixture:
DNS N=8192 (FP64, Re=1e6, K0=8, seed 0, fixed dt = the initial CFL step, 40000 steps =
5000 coarse steps of the N=1024 run), spectrally truncated to the N=1024 grid: the shell
energy spectrum at t_end (E_dns, shells 0..1023), the enstrophy at 5 equally spaced coarse
steps (z_dns), dt_f, t_end and the DNS's GPU seconds (t_dns), under the key that
optimizec_amd.closure_eval checks (--dns-cache).  The FP64 GPU DNS is the reference: the
oracle pins the same path at N=8192 (tests/test_gpu_multipass.py).

Run once on an MI355X from the repository root (about 10 minutes of GPU time):

    python tests/golden/make_closure_reference.py [out.npz]

(default out: tests/golden/closure_dns_N8192_Re1e6_K08_seed0.npz).  It also prints the
closure evaluation against the new reference as one JSON line.
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from optimizec_amd import closure_eval  # noqa: E402

DEFAULT = os.path.join(ROOT, "tests", "golden", "closure_dns_N8192_Re1e6_K08_seed0.npz")

if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    if os.path.exists(out):
        os.remove(out)
    r = closure_eval.evaluate(coarse=1024, ratio=8, Re=1e6, K0=8.0, seed=0, coarse_steps=5000, ensemble="config",
                              train_iters=4000, log=sys.stderr, dns_cache=out)
    print(json.dumps(r), flush=True)
