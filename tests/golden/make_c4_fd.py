"""Writes tests/golden/c4_fd.json: FP64 central finite differences of the
centroid loss on the differentiable scene at full size (65,536 particles,
64^3 grid, 64 substeps), for 32 seeded parameters -- 16 components of the
initial velocity and 16 per-particle Young's moduli.

The forward is the oracle's CPUDIFF restatement (mpm_cpu_diff.rs
forward_with_tape / step_once_record, f64); the loss is
J = |COM(x_T) - target|^2 (mpm_adjoint.rs:516-556).  Each derivative is
(J(p + h) - J(p - h)) / 2h, taken at two step sizes h and h/4 so that the
fixture also records its own truncation error (Richardson difference).

Beside the differences it records the oracle's adjoint (mpm_adjoint.rs
backward restated) of the same parameters twice: from the workload's F0 = I,
where the SVD adjoint's Tikhonov term (dense3.rs:295) zeroes the couplings of
the exactly repeated singular values, and from F0 = I + 1e-9 N(0, 1)
(F0_NOISE_SEED), off that set.  Their difference quantifies the F0 = I
caveat of any adjoint compared with these differences.

Run from the repository root after building the oracle:
    python tests/golden/make_c4_fd.py
The 129 forwards run on all host threads (the oracle releases the GIL).
"""
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as orc  # noqa: E402
from molla_amd import workloads  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_fd.json")
SEED = 20261203
F0_NOISE_SEED = 5


def f0_noise(n):
    return np.eye(3)[None] + 1e-9 * np.random.default_rng(F0_NOISE_SEED).standard_normal((n, 3, 3))


def oracle_grads(cfg, d, meta, F0=None):
    """Oracle adjoint of the centroid loss: (dJ/dv0 (n, 3), dJ/dE (n,))."""
    n = len(d["x"])
    st = orc.State(d["x"], d["v"], F=F0, mass=np.full(n, cfg.particle_mass), vol=np.full(n, cfg.particle_volume),
                   E=d["E"])
    st0 = st.copy()
    tape = orc.diff_forward(cfg, st, meta["dt"], meta["substeps"])
    r = st.x.mean(axis=0) - np.asarray(meta["target"])
    g = orc.diff_backward(cfg, st0, meta["dt"], meta["substeps"], tape, np.tile((2.0 / n) * r, (n, 1)))
    return g["v"], g["E"]


def loss(cfg, d, meta, v, E):
    n = len(d["x"])
    st = orc.State(d["x"], v, mass=np.full(n, cfg.particle_mass), vol=np.full(n, cfg.particle_volume), E=E)
    orc.diff_forward(cfg, st, meta["dt"], meta["substeps"], record=False)
    r = st.x.mean(axis=0) - np.asarray(meta["target"])
    return float(r @ r)


def main():
    cfg, d, meta = workloads.diff_scene()
    n = len(d["x"])
    rng = np.random.default_rng(SEED)
    vp = rng.choice(n, 16, replace=False)
    va = rng.integers(0, 3, 16)
    ep = rng.choice(n, 16, replace=False)
    v0 = d["v"].astype(np.float64)
    E0 = d["E"].astype(np.float64)
    params = [("v", int(p), int(a)) for p, a in zip(vp, va)] + [("E", int(p), 0) for p in ep]

    def step_of(kind, p, a):
        return 1e-3 if kind == "v" else 1e-3 * float(E0[p])

    jobs = [("adjoint", 0), ("adjoint", 1), (None, 0.0)]
    for i, (kind, p, a) in enumerate(params):
        h = step_of(kind, p, a)
        for s in (h, -h, h / 4, -h / 4):
            jobs.append((i, s))

    def run(job):
        i, s = job
        if i == "adjoint":
            return oracle_grads(cfg, d, meta, None if s == 0 else f0_noise(n))
        v, E = v0.copy(), E0.copy()
        if i is not None:
            kind, p, a = params[i]
            if kind == "v":
                v[p, a] += s
            else:
                E[p] += s
        return loss(cfg, d, meta, v, E)

    t0 = time.time()
    with ThreadPoolExecutor(max_workers=os.cpu_count() or 8) as ex:
        vals = list(ex.map(run, jobs))
    (gv_i, ge_i), (gv_p, ge_p), J0 = vals[0], vals[1], vals[2]
    rows = []
    for i, (kind, p, a) in enumerate(params):
        jp, jm, jq, jr = vals[3 + 4 * i: 7 + 4 * i]
        h = step_of(kind, p, a)
        rows.append({"param": kind, "particle": p, "axis": a, "h": h,
                     "fd": (jp - jm) / (2 * h), "fd_quarter": (jq - jr) / (h / 2),
                     "oracle_adjoint_F0_I": float(gv_i[p, a] if kind == "v" else ge_i[p]),
                     "oracle_adjoint_F0_noise": float(gv_p[p, a] if kind == "v" else ge_p[p])})
    out = {"workload": "diff_scene", "particles": n, "grid": meta["grid"], "substeps": meta["substeps"],
           "dt": meta["dt"], "target": meta["target"], "seed": SEED, "f0_noise_seed": F0_NOISE_SEED, "loss": J0,
           "generator": "tests/golden/make_c4_fd.py (oracle CPUDIFF forward, f64)", "rows": rows}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}: {len(rows)} derivatives, loss {J0:.6e}, {time.time() - t0:.0f} s")
    for r in rows:
        fd = r["fd"]
        print(f"{r['param']} {r['particle']:6d} fd {fd: .6e} (richardson {abs(fd - r['fd_quarter']) / abs(fd):.1e}) "
              f"adjoint F0=I {abs(r['oracle_adjoint_F0_I'] - fd) / abs(fd):.1e} "
              f"F0=I+noise {abs(r['oracle_adjoint_F0_noise'] - fd) / abs(fd):.1e}")


if __name__ == "__main__":
    main()
