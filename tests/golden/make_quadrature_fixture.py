"""Extract the Gauss-Legendre known-answer tables held by the reference's
test (Sources/Tests/Serial/test_quadratures/test_quadratures.f90:40-500) into
gauss_legendre_tables.json.  Data only: abscissae and weights for 1..20 points.
Run in the build container, where the reference is mounted read-only."""
import json
import os
import re

SRC = "/root/reference/Sources/Tests/Serial/test_quadratures/test_quadratures.f90"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gauss_legendre_tables.json")


def main():
    text = open(SRC).read()
    tables = {}
    cur = None
    for line in text.splitlines():
        m = re.search(r"nlocs\s*==\s*(\d+)", line)
        if m and ("if" in line):
            cur = int(m.group(1))
            tables[cur] = {"x": {}, "w": {}}
            continue
        m = re.search(r"(posgl|weigl)\s*\(\s*(\d+)\s*\)\s*=\s*([-+0-9.eEdD]+)_rp", line)
        if m and cur is not None:
            kind = "x" if m.group(1) == "posgl" else "w"
            tables[cur][kind][int(m.group(2))] = float(m.group(3).replace("d", "e").replace("D", "e"))
    out = {}
    for n, t in sorted(tables.items()):
        xs = [t["x"].get(i, 0.0) for i in range(1, n + 1)]
        ws = [t["w"].get(i, 0.0) for i in range(1, n + 1)]
        out[str(n)] = {"x": xs, "w": ws}
    json.dump({"source": "Sources/Tests/Serial/test_quadratures/test_quadratures.f90", "tolerance": 1e-10,
               "tables": out}, open(OUT, "w"), indent=1)
    print("wrote", OUT, "with", len(out), "rules")


if __name__ == "__main__":
    main()
