"""Input and output files of the Fortran driver, in the reference's formats.

control.prm  positional parameters (main.f90:222-277)
input.xyz    two header lines, then 'P x y z id' (electrons) and 'O x y z id'
             (ions); a restart file adds velocities (main.f90:298-380)
ener###.dat  9 columns ES14.6: t(ns), T_e, T_i, T_e,rad, T_i,rad, T_e,cir,
             T_i,cir, Epot, Etot (print.f90:90-93)
res###.xyz   restart frame (sample.f90:69-94)
"""
import numpy as np

from .engine import PT_DTYPE, XM_E, XM_I


def write_control(path, ne, ni, te=10.0, ti=1e-3, tmax=0.05, trest=1.0, tdump=1.0, tsamp=0.005, dt=0.005,
                  rs=6.6493e-3, theta=0.5, rest="OFF", dx=(0.01, 0.01, 0.01), xm=(XM_E, XM_I)):
    """Times in ns, lengths in um, masses in amu."""
    lines = [
        "#control", "#charge", "-1.0", "1.0", "#mass", repr(xm[0]), repr(xm[1]),
        "#number", str(ne), str(ni), "#temper", repr(te), repr(ti),
        "#time", "%r %r %r %r %r" % (tmax, trest, tdump, tsamp, dt),
        "#kelbg", "ON %r" % rs, "#restart", rest, "#rdf", "%r %r %r" % dx, "#tree", repr(theta),
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_xyz(path, qel, qion, velocities=False):
    with open(path, "w") as f:
        f.write("%d\n" % (len(qel) + len(qion)))
        f.write("frame = 0 energy = 0\n")
        for tag, q in (("P", qel), ("O", qion)):
            for p in q:
                if velocities:
                    f.write("%s %.17g %.17g %.17g %.17g %.17g %.17g %d\n" % ((tag,) + tuple(p["xx"]) +
                                                                              tuple(p["xv"]) + (p["id"],)))
                else:
                    f.write("%s %.17g %.17g %.17g %d\n" % ((tag,) + tuple(p["xx"]) + (p["id"],)))


def read_ener(path):
    return np.loadtxt(path, comments="#", ndmin=2)


def read_rdf(path):
    """rdf###.dat (sample.f90:34-38): columns R_ei, RDF_ei, R_ii, RDF_ii (counts per step)."""
    return np.loadtxt(path, comments="#", ndmin=2)


def read_res(path):
    """(qel, qion) from a res###.xyz restart frame."""
    rows = [ln.split() for ln in open(path).read().splitlines()[2:] if ln.strip()]
    e = [r for r in rows if r[0] == "P"]
    i = [r for r in rows if r[0] == "O"]
    out = []
    for rr in (e, i):
        q = np.zeros(len(rr), PT_DTYPE)
        a = np.array([[float(v) for v in r[1:7]] for r in rr]) if rr else np.zeros((0, 6))
        q["xx"] = a[:, :3]
        q["xv"] = a[:, 3:6]
        q["id"] = [int(r[7]) for r in rr]
        out.append(q)
    return out[0], out[1]
