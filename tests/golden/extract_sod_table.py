"""Regenerate sod_exact_t0245.txt from the reference's regression test (data only).

Runs only where the reference checkout is mounted (/root/reference)."""
import re
import sys

SRC = "/root/reference/tst/regression/test_suites/riemann_hydro/riemann_hydro.py"


def main(out="sod_exact_t0245.txt"):
    body = re.search(r'sod_solution = """(.*?)"""', open(SRC).read(), re.S).group(1)
    rows = [list(map(float, l.split())) for l in body.strip().splitlines()]
    with open(out, "w") as f:
        f.write("# Exact Sod solution (Toro test 1) at t=0.245, gamma=1.4, from the reference's\n")
        f.write("# tst/regression/test_suites/riemann_hydro/riemann_hydro.py:192-269 (x in that table is\n")
        f.write("# shifted by +1.0 at :271-272 for the x in [0,2], x_discont=1 run).\n")
        f.write("# columns: x_unshifted rho v p\n")
        for r in rows:
            f.write(" ".join(f"{v:.16e}" for v in r) + "\n")


if __name__ == "__main__":
    main(*sys.argv[1:])
