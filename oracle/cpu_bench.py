"""CPU baseline: the oracle's SIMPLE iteration timed on one host core.

TEST/MEASUREMENT INFRASTRUCTURE: bench.py runs this in a child process as its
cpu_baseline leg.  The oracle (oracle/mc_oracle.c, single-threaded C, the
reference CPU algorithm with an aggregation-AMG preconditioned CG pressure
solve in place of PETSc GAMG+FGMRES) runs the same workload as the GPU bench
(box mesh modifier m, simple mode, pressure rel. residual 1e-8, BiCGStab
rtol 1e-5, momentum and scalar BiCGStab started from the current fields as
the GPU forms start them).  Sample: one untimed SIMPLE iteration (it performs
the one-off set_up_field and the AMG aggregation), then --iters timed SIMPLE
iterations (default 3), run while no GPU work is timed.  Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import pyoracle as po  # noqa: E402


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--m", type=int, default=128)
    ap.add_argument("--iters", type=int, default=3)
    args = ap.parse_args()
    t_setup = time.perf_counter()
    o = po.Oracle(args.m, mode=po.MODE_SIMPLE, mom_rtol=1e-5, p_rtol=1e-8, p_solver=1, warm_start=1)
    o.simple_iteration(0)
    t0 = time.perf_counter()
    for i in range(args.iters):
        o.simple_iteration((1 + i) % 3)
    t1 = time.perf_counter()
    rate = args.iters / (t1 - t0)
    print(json.dumps({
        "value": rate,
        "unit": "SIMPLE iter/s",
        "cores": 1,
        "kind": "port",
        "sample": "CPU oracle (C, one thread), box modifier %d (%d cells), simple mode, AMG-PCG to 1e-8, "
                  "BiCGStab 1e-5 from the current fields; %d timed SIMPLE iterations after one untimed; "
                  "host %s, nproc %d; setup+warmup %.1f s"
                  % (args.m, o.N, args.iters, cpu_model(), os.cpu_count() or 0, t0 - t_setup),
        "seconds": t1 - t0,
        "pressure_iters": int(o.s.stats.p_iters),
    }))
    o.close()


if __name__ == "__main__":
    main()
