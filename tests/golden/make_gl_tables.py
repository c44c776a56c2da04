"""Extract the Gauss-Legendre node/weight tables GLQuadrature<N> of the reference
(DiFfRG/include/DiFfRG/common/quadrature/gauss_legendre.hh) into gl_tables.json.

The values are kept as decimal strings (30+ significant digits) so the tests can
compare in extended precision.  Run from this container, where the reference is
mounted; the JSON it writes is the committed fixture.
"""
import json
import os
import re
import sys

SRC = "/root/reference/DiFfRG/include/DiFfRG/common/quadrature/gauss_legendre.hh"


def main(out_path: str) -> None:
    text = open(SRC).read()
    tables = {}
    for m in re.finditer(r"struct GLQuadrature<(\d+), ctype>\s*\{(.*?)\n  \};", text, re.S):
        n = int(m.group(1))
        body = m.group(2)
        xs = dict((int(i), v) for i, v in re.findall(r"\bx\[(\d+)\]\s*=\s*([-+0-9.eE]+);", body))
        ws = dict((int(i), v) for i, v in re.findall(r"\bw\[(\d+)\]\s*=\s*([-+0-9.eE]+);", body))
        assert len(xs) == n and len(ws) == n, n
        tables[str(n)] = {"x": [xs[i] for i in range(n)], "w": [ws[i] for i in range(n)]}
    json.dump({"source": "DiFfRG/include/DiFfRG/common/quadrature/gauss_legendre.hh",
               "interval": "[0,1]", "orders": sorted(int(k) for k in tables), "tables": tables},
              open(out_path, "w"), indent=0)
    print("wrote", out_path, sorted(int(k) for k in tables))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "gl_tables.json"))
