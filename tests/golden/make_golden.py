"""Generate the golden vectors under tests/golden/ from the reference's own
tree code (tree.f90 / data.f90 / vverlet.f90 compiled from /root/reference
into oracle/_ref/ref_harness by oracle/Makefile).  Run in the build container:

    make -C oracle all ref && python tests/golden/make_golden.py

Fixtures hold inputs and the reference's outputs only.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import tree_amd as ta  # noqa: E402  (host-only generator)
from oracle import pyoracle  # noqa: E402

RS10 = 6.6493e-3
DT = 0.005 / ta.TPS
RDF_DX = 0.01  # the harness's RDF bin width (dx_ei = dx_ii), um


def tree_arrays(t):
    """(depth, ileaf, pt, next) and the 17 doubles; pt of a twig (0 once a leaf
    turns into a twig, uninitialised at the root) is normalised to 0."""
    ti = t["i"].copy()
    ti[:, 2] = np.where(ti[:, 1] != 0, ti[:, 2], 0)
    return ti, t["d"].copy()


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def small_case(name, xe, xi, theta):
    z = np.zeros_like
    buf = pyoracle.run_ref(1, xe, z(xe), xi, z(xi), theta, RS10, (ta.XM_E, ta.XM_I), DT, threads=1)
    trees, fe, fi, ep = pyoracle.parse_ref_mode1(buf, len(xe), len(xi))
    rdf_ei, rdf_ii = pyoracle.parse_ref_rdf(buf, len(xe), len(xi))
    out = dict(xe=xe, xi=xi, theta=theta, rs=RS10, fe=fe, fi=fi, epot=ep, rdf_dx=RDF_DX, rdf_ei=rdf_ei,
               rdf_ii=rdf_ii)
    for sp, t in enumerate(trees):
        out["tree%d_int" % sp], out["tree%d_dbl" % sp] = tree_arrays(t)
    np.savez_compressed(os.path.join(HERE, name), **out)
    print(name, len(trees[0]), len(trees[1]), ep)


def main():
    if not pyoracle.ref_available():
        sys.exit("oracle/_ref/ref_harness missing: make -C oracle ref")
    rng = np.random.default_rng(2024)
    # 1. small Gaussian case: full node tables
    small_case("ref_n96.npz", rng.normal(size=(96, 3)) * 100.0, rng.normal(size=(90, 3)) * 100.0, 0.5)
    # 2. lattice with ties on cell centres, coarse theta
    g = np.arange(-2, 3, dtype=np.float64) * 10.0
    lat = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
    small_case("ref_lattice.npz", lat, lat * 0.5 + 1.25, 0.7)
    # 3. C1: 4096 + 4096, sigma 100 um, seed 1 from the engine's generator
    qel, qion = ta.gen_plasma(1, 4096, 4096, 100.0)
    xe, xi = qel["xx"].copy(), qion["xx"].copy()
    z = np.zeros_like
    buf = pyoracle.run_ref(1, xe, z(xe), xi, z(xi), 0.5, RS10, (ta.XM_E, ta.XM_I), DT, threads=1)
    trees, fe, fi, ep = pyoracle.parse_ref_mode1(buf, 4096, 4096)
    rdf_ei, rdf_ii = pyoracle.parse_ref_rdf(buf, 4096, 4096)
    meta = dict(pos_digest=digest(np.r_[xe, xi]), fe=fe, fi=fi, epot=ep, theta=0.5, rs=RS10,
                nnodes=np.array([len(trees[0]), len(trees[1])]), rdf_dx=RDF_DX, rdf_ei=rdf_ei, rdf_ii=rdf_ii)
    for sp, t in enumerate(trees):
        ti, td = tree_arrays(t)
        meta["tree%d_int_digest" % sp] = digest(ti)
        meta["tree%d_dbl_digest" % sp] = digest(td)
    np.savez_compressed(os.path.join(HERE, "ref_c1.npz"), **meta)
    print("ref_c1.npz", meta["nnodes"], ep)
    # 4. trajectory: 200 + 200 particles, 10 velocity-Verlet steps with the reference's integrator
    qel, qion = ta.gen_plasma(5, 200, 200, 30.0)
    xe, ve, xi, vi = qel["xx"].copy(), qel["xv"].copy(), qion["xx"].copy(), qion["xv"].copy()
    buf = pyoracle.run_ref(2, xe, ve, xi, vi, 0.5, RS10, (ta.XM_E, ta.XM_I), DT, nsteps=10)
    steps, e, i = pyoracle.parse_ref_mode2(buf, 200, 200, 10)
    np.savez_compressed(os.path.join(HERE, "ref_traj.npz"), xe=xe, ve=ve, xi=xi, vi=vi, theta=0.5, rs=RS10,
                        dt=DT, steps=steps, e=e, i=i)
    print("ref_traj.npz", steps[-1])


if __name__ == "__main__":
    main()
