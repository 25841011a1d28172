"""Write tests/golden/config1_rk4.npz: the FP64 oracle's run of the 128x32x1
LLG-limit film (hotspin_amd.workloads.thin_film, seed 0), fixed RK4 at
dt = 1e-13 s for 10,000 steps: <m>(t) every 100 steps and the final m."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hotspin_amd import workloads as W  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def main(steps=10000, every=100, dt=1e-13):
    N, cell, kw, m = W.thin_film(0)
    o = po.LLB(N, cell, **kw)
    trace = [m.mean(axis=1)]
    t0 = time.time()
    for i in range(steps):
        o.step("rk4", m, dt, i)
        if (i + 1) % every == 0:
            mn = m / np.linalg.norm(m, axis=0)
            trace.append(mn.mean(axis=1))
    print("oracle: %d RK4 steps in %.1f s" % (steps, time.time() - t0))
    np.savez_compressed(os.path.join(HERE, "config1_rk4.npz"), m_final=m, avg_trace=np.array(trace),
                        steps=steps, every=every, dt=dt)


if __name__ == "__main__":
    main()
