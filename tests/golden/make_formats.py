"""Writes tests/golden/formats.txt: lines in the output formats of the
reference's ener###.dat (TEMPORAL, sample.f90:90-93: 9(ES14.6, 1x)),
res###.xyz (Restart, print.f90:73-106: "P  " / "O  ", 6(ES14.6, 1X), I8) and
rdf###.dat (RDF, sample.f90:36-43: 5(ES14.6, 1X), four values written), as
amdflang's formatted output writes them for fixed values (signs, zero, a
three-digit exponent, rounding at the sixth decimal).

Run from anywhere: python tests/golden/make_formats.py  (needs amdflang).
The values and the lines go to formats.txt: '#V' lines hold the values
(repr), the next line the formatted record.
"""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")

VALUES = [1.0, -2.5e-3, 123456.789, -9.87654321e10, 0.0, 1.0e-120, -7.77777777e99, 9.9999995, 0.33]

SRC = """program fmt
  implicit none
  real*8 :: v(9)
  v = (/ %s /)
  write(*, 500) v
  write(*, 80) v(1:6), 12
  write(*, 90) v(4:9), 12345678
  write(*, 200) v(1), v(2), v(3), v(4)
500 format(9(ES14.6, 1x))
80 format("P  ", 6(ES14.6, 1X), I8)
90 format("O  ", 6(ES14.6, 1X), I8)
200 format(5(ES14.6, 1X))
end program
"""


def main():
    lit = ", ".join(("%.17e" % v).replace("e", "d") for v in VALUES)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "fmt.f90"), "w") as f:
            f.write(SRC % lit)
        subprocess.run([FLANG, "fmt.f90", "-o", "fmt"], cwd=d, check=True)
        out = subprocess.run([os.path.join(d, "fmt")], capture_output=True, text=True, check=True).stdout
    recs = out.splitlines()
    vals = [(VALUES, None), (VALUES[:6], 12), (VALUES[3:9], 12345678), (VALUES[:4], None)]
    kinds = ["ener", "res_P", "res_O", "rdf"]
    with open(os.path.join(HERE, "formats.txt"), "w") as f:
        for kind, (v, i), rec in zip(kinds, vals, recs):
            f.write("#V %s %s %s\n" % (kind, " ".join(repr(x) for x in v), "" if i is None else i))
            f.write(rec + "\n")


if __name__ == "__main__":
    main()
