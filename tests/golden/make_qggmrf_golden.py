"""Writes tests/golden/qggmrf_golden.json from the golden vectors that the
reference's own test holds (mbirjax tests/test_qggmrf.py:83-134: grad_ref,
hess_ref, the 3x3x3 input flat_recon and the parameters p, q, T, sigma_x with
b = six equal weights).  The reference file is read as text; nothing in it is
imported or executed.  Run once in a container that has /root/reference.
"""
import json
import os
import re
import sys

SRC = "/root/reference/tests/test_qggmrf.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qggmrf_golden.json")


def array_after(text, name):
    i = text.index(name)
    j = text.index("np.array(", i) + len("np.array(")
    depth, k = 0, j
    while True:
        c = text[k]
        if c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                break
        k += 1
    body = text[j:k + 1]
    rows = re.findall(r"\[([^\[\]]+)\]", body)
    return [[float(v) for v in r.split(",") if v.strip()] for r in rows]


def scalar(text, name):
    m = re.search(r"\n\s*%s = ([0-9.eE+-]+)" % name, text)
    return float(m.group(1))


def main():
    text = open(SRC).read()
    seg = text[text.index("def test_gradient_and_hessian"):text.index("def test_loss_and_gradient")]
    golden = {
        "source": "mbirjax 0.6.7 tests/test_qggmrf.py:83-134",
        "recon_shape": [3, 3, 3],
        "p": scalar(seg, "p"), "q": scalar(seg, "q"), "T": scalar(seg, "T"), "sigma_x": scalar(seg, "sigma_x"),
        "b": [1.0 / 6] * 6,
        "flat_recon": array_after(seg, "flat_recon ="),
        "grad_ref": array_after(seg, "grad_ref ="),
        "hess_ref": array_after(seg, "hess_ref ="),
    }
    assert len(golden["flat_recon"]) == 9 and len(golden["grad_ref"]) == 9 and len(golden["hess_ref"]) == 9
    with open(OUT, "w") as f:
        json.dump(golden, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
