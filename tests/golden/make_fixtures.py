"""Generate the golden fixtures of tests/golden/ from the reference tree.

Run in the build container (the reference is not on the GPU box):
    python tests/golden/make_fixtures.py [/root/reference]

Writes
  eam_constants.json  every `double NAME = NUMBER;` of potential.cpp, per EAM
                      function, in order of appearance, plus the literal scale
                      divisor (`x /= S;`) and switch radius (`else if (x>R)`)
                      of each body -- numbers read from the reference text.
  quadrature_ref.json the nodes printed by oracle/_ref/quadrature_ref, the
                      reference's own gaussianQuadrature(6, QP) (quadrature.cpp)
                      compiled from its source, as hex floats.
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def eam_constants(ref):
    txt = open(os.path.join(ref, "potential.cpp")).read()
    heads = [(m.start(), m.group(1)) for m in re.finditer(r"double\s+EAM::(\w+)\s*\(", txt)]
    out = {}
    for k, (pos, name) in enumerate(heads):
        if name.endswith("_deriv"):
            continue
        end = heads[k + 1][0] if k + 1 < len(heads) else len(txt)
        body = txt[pos:end]
        consts = [[n, float(v)] for n, v in re.findall(r"double\s+(\w+)\s*=\s*([-+0-9.eE]+)\s*;", body)]
        scale = re.findall(r"\b(?:r|rho)\s*/=\s*([0-9.]+)\s*;", body)
        switch = re.findall(r"else\s+if\s*\(\s*r(?:\*5\.0)?\s*>\s*([0-9.]+)\s*\)", body)
        out[name] = {"constants": consts,
                     "scale": float(scale[0]) if scale else None,
                     "switch": float(switch[0]) if switch else None}
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "eam_constants.json"), "w") as f:
        json.dump(eam_constants(ref), f, indent=1)
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref", "REF=" + ref], check=True)
    out = subprocess.run([os.path.join(REPO, "oracle", "_ref", "quadrature_ref")], capture_output=True,
                         text=True, check=True).stdout.split("\n")
    k = int(out[0])
    nodes = [line.split() for line in out[1:1 + k]]
    with open(os.path.join(HERE, "quadrature_ref.json"), "w") as f:
        json.dump({"source": "quadrature.cpp:43-75 gaussianQuadrature(6, QP)", "nodes_hex": nodes}, f, indent=1)
    print("fixtures written")


if __name__ == "__main__":
    main()
