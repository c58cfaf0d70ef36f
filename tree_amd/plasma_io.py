"""Input and output files of the Fortran driver, in the reference's formats.

control.prm  positional parameters (main.f90:222-277)
input.xyz    two header lines, then 'P x y z id' (electrons) and 'O x y z id'
             (ions); a restart file adds velocities (main.f90:298-380)
ener###.dat  9 columns ES14.6: t(ns), T_e, T_i, T_e,rad, T_i,rad, T_e,cir,
             T_i,cir, Epot, Etot (TEMPORAL, sample.f90:90-93)
res###.xyz   restart frame (Restart, print.f90:67-106)
rdf###.dat   R_ei, RDF_ei, R_ii, RDF_ii (RDF, sample.f90:32-43)

The format_* helpers write one record as those FORMAT statements do (ES14.6
fields separated by 1X; a three-digit exponent drops the 'E', as Fortran's
ES editing does), so tests can check the driver's files line by line.
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


def es14_6(x):
    """One ES14.6 field (Fortran scientific editing, width 14, 6 decimals)."""
    m, e = ("%.6E" % x).split("E")
    ex = int(e)
    s = "%sE%+03d" % (m, ex) if abs(ex) < 100 else "%s%+04d" % (m, ex)
    return "%14s" % s


def format_ener(row):
    """A TEMPORAL record: 9(ES14.6, 1x)."""
    return " ".join(es14_6(v) for v in row)


def format_res(tag, xx, xv, pid):
    """A Restart record: "P  " or "O  ", 6(ES14.6, 1X), I8."""
    return tag + "  " + " ".join(es14_6(v) for v in list(xx) + list(xv)) + " " + "%8d" % pid


def format_rdf(vals):
    """An RDF record: 5(ES14.6, 1X) with the four values written."""
    return " ".join(es14_6(v) for v in vals)


def fortran_float(tok):
    """A real as Fortran writes it, including the E-less three-digit exponent (1.000000-120)."""
    t = tok.replace("D", "E").replace("d", "E")
    if "E" not in t and "e" not in t:
        k = max(t.rfind("+"), t.rfind("-"))
        if k > 0:
            t = t[:k] + "E" + t[k:]
    return float(t)


def _read_table(path):
    rows = [[fortran_float(v) for v in ln.split()] for ln in open(path).read().splitlines()
            if ln.strip() and not ln.lstrip().startswith("#")]
    return np.array(rows, dtype=np.float64).reshape(len(rows), -1)


def read_ener(path):
    return _read_table(path)


def read_rdf(path):
    """rdf###.dat (sample.f90:32-43): columns R_ei, RDF_ei, R_ii, RDF_ii (counts per step)."""
    return _read_table(path)


def read_res(path):
    """(qel, qion) from a res###.xyz restart frame."""
    rows = [ln.split() for ln in open(path).read().splitlines()[2:] if ln.strip()]
    e = [r for r in rows if r[0] == "P"]
    i = [r for r in rows if r[0] == "O"]
    out = []
    for rr in (e, i):
        q = np.zeros(len(rr), PT_DTYPE)
        a = np.array([[fortran_float(v) for v in r[1:7]] for r in rr]) if rr else np.zeros((0, 6))
        q["xx"] = a[:, :3]
        q["xv"] = a[:, 3:6]
        q["id"] = [int(r[7]) for r in rr]
        out.append(q)
    return out[0], out[1]
