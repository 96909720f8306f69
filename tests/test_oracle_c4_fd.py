"""The oracle's adjoint (mpm_adjoint.rs backward restated, f64) on the full
differentiable scene (65,536 particles, 64^3, 64 substeps, centroid loss,
mpm_adjoint.rs:516-556) against FP64 central finite differences of the
oracle's CPUDIFF forward on 32 seeded parameters (16 components of v0, 16
per-particle E), written by tests/golden/make_c4_fd.py.

The adjoint is taken from F0 = I + 1e-9 N(0, 1): at exactly F0 = I the
reference's SVD adjoint zeroes the couplings of repeated singular values
(Tikhonov term, dense3.rs:295), which moves dJ/dv0 by up to 1.4e-3 of itself
on this scene (the fixture records both adjoints beside the differences).

Bound per parameter: |adjoint - fd| <= 1e-4 |fd| + 4 |fd - fd(h/4)|, the last
term being the differences' own truncation and rounding error (Richardson);
parameters whose difference is pure rounding (an E far from any contact has
|dJ/dE| ~ 1e-19) pass through that term, and at least 20 of the 32 must be
resolved to 1 % by it so the check is not vacuous."""
import json
import os

import numpy as np

import oracle as orc
from molla_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture():
    with open(os.path.join(HERE, "golden", "c4_fd.json")) as f:
        return json.load(f)


def test_oracle_adjoint_matches_c4_finite_differences():
    fx = _fixture()
    cfg, d, meta = workloads.diff_scene()
    n = len(d["x"])
    assert n == fx["particles"] and meta["substeps"] == fx["substeps"]
    F0 = np.eye(3)[None] + 1e-9 * np.random.default_rng(fx["f0_noise_seed"]).standard_normal((n, 3, 3))
    st = orc.State(d["x"], d["v"], F=F0, mass=np.full(n, cfg.particle_mass), vol=np.full(n, cfg.particle_volume),
                   E=d["E"])
    st0 = st.copy()
    tape = orc.diff_forward(cfg, st, meta["dt"], meta["substeps"])
    r = st.x.mean(axis=0) - np.asarray(meta["target"])
    assert abs(float(r @ r) - fx["loss"]) <= 1e-9 * fx["loss"]
    g = orc.diff_backward(cfg, st0, meta["dt"], meta["substeps"], tape, np.tile((2.0 / n) * r, (n, 1)))
    resolved = 0
    for row in fx["rows"]:
        adj = g["v"][row["particle"], row["axis"]] if row["param"] == "v" else g["E"][row["particle"]]
        fd, noise = row["fd"], abs(row["fd"] - row["fd_quarter"])
        resolved += noise <= 1e-2 * abs(fd)
        assert abs(adj - fd) <= 1e-4 * abs(fd) + 4 * noise, (row, adj)
    assert resolved >= 20, resolved
