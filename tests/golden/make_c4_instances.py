"""Draw the 512-instance parameter table of the batched workload (C4) and write
``c4_instances.json`` beside this script.

numpy ``default_rng(0)``; 256 telegraph then 256 toggle instances, parameters
log-uniform over the stated ranges.
* telegraph: k_on, k_off in [1e-3, 1], k_r in [1, 100], gamma in [0.01, 1]; the mRNA
  bound is mean + 10 sd of the stationary law (mean = k_r k_on / ((k_on + k_off) gamma),
  Fano = 1 + k_r k_off / ((k_on + k_off)(k_on + k_off + gamma))), at least 99, so that
  every space has 200 states or more.
* toggle: alpha in [10, 100], K in [5, 50], gamma in [0.01, 1]; basal rate 1 and Hill
  exponent 2.5 as in the fixed toggle; per-species bound m + 10 sqrt(m) with the Poisson
  mean m = (alpha + 1) / gamma. A draw whose box (bound + 1)^2 exceeds 200,000 states is
  rejected and drawn again (all three parameters), so every instance still covers its
  mean + 10 sd; the rejections fall on small gamma and large alpha.
"""
import json
import math
from pathlib import Path

import numpy as np

MAX_STATES = 200_000


def draw():
    rng = np.random.default_rng(0)
    lu = lambda a, b: float(10 ** rng.uniform(math.log10(a), math.log10(b)))
    out = []
    for _ in range(256):
        kon, koff, kr, g = lu(1e-3, 1), lu(1e-3, 1), lu(1, 100), lu(0.01, 1)
        mean = kr * kon / ((kon + koff) * g)
        fano = 1 + kr * koff / ((kon + koff) * (kon + koff + g))
        nmax = max(99, int(math.ceil(mean + 10 * math.sqrt(mean * fano))))
        out.append({"kind": "telegraph", "theta": [kon, koff, kr, g], "nmax": nmax, "states": 2 * (nmax + 1)})
    rejected = 0
    while len(out) < 512:
        alpha, K, g = lu(10, 100), lu(5, 50), lu(0.01, 1)
        m = (alpha + 1.0) / g
        nmax = int(math.ceil(m + 10 * math.sqrt(m)))
        if (nmax + 1) ** 2 > MAX_STATES:
            rejected += 1
            continue
        out.append({"kind": "toggle", "theta": [alpha, K, 2.5, 1.0, g], "nmax": nmax, "states": (nmax + 1) ** 2})
    return out, rejected


if __name__ == "__main__":
    table, rejected = draw()
    (Path(__file__).resolve().parent / "c4_instances.json").write_text(json.dumps(table, indent=0))
    sizes = [t["states"] for t in table]
    print(len(table), "instances; states", min(sizes), "..", max(sizes), "total", sum(sizes),
          "; toggle draws rejected:", rejected)
