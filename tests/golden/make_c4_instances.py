"""Draw the 512-instance parameter table of the batched workload (C4) and write
``c4_instances.json`` beside this script.

numpy ``default_rng(0)``; 256 telegraph then 256 toggle instances.
* telegraph: k_on, k_off log-uniform in [1e-3, 1], k_r in [1, 100], gamma in [0.01, 1];
  the mRNA bound is mean + 10 sd of the stationary law (mean = k_r k_on / ((k_on + k_off)
  gamma), Fano = 1 + k_r k_off / ((k_on + k_off)(k_on + k_off + gamma))), at least 99 so
  that every space has 200 states or more.
* toggle: alpha log-uniform in [10, 100], K in [5, 50]; basal rate 1, Hill exponent 2.5
  and degradation 1 as in the fixed toggle; per-species bound m + 10 sqrt(m) with the
  Poisson mean m = alpha + 1, at least 14.
"""
import json
import math
from pathlib import Path

import numpy as np


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
    for _ in range(256):
        alpha, K = lu(10, 100), lu(5, 50)
        m = alpha + 1.0
        nmax = max(14, int(math.ceil(m + 10 * math.sqrt(m))))
        out.append({"kind": "toggle", "theta": [alpha, K, 2.5, 1.0, 1.0], "nmax": nmax, "states": (nmax + 1) ** 2})
    return out


if __name__ == "__main__":
    table = draw()
    (Path(__file__).resolve().parent / "c4_instances.json").write_text(json.dumps(table, indent=0))
    sizes = [t["states"] for t in table]
    print(len(table), "instances; states", min(sizes), "..", max(sizes), "total", sum(sizes))
