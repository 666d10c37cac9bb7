"""Write tests/golden/sampling_vectors.json: the seeded boson-sampling cases of
the reference's tests (tests/_simulators/passive/test_measurements.py:237-563),
transcribed as data: the 5-mode interferometer, the input occupation, the
seed, the shot count and the expected sample list of each case.

The literals are read from the reference's test file with ``ast`` (nothing is
imported or run).  Run once in a container that has the reference checked
out at /root/reference:

    python tests/golden/make_sampling_vectors.py
"""

import ast
import json
import os

SRC = "/root/reference/tests/_simulators/passive/test_measurements.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sampling_vectors.json")

# test name -> how the reference reaches the sampler for it
CASES = {
    "test_boson_sampling_seeded": "ideal",
    "test_LossyInterferometer_boson_sampling_seeded": "lossy_nonuniform",
    "test_LossyInterferometer_boson_sampling_uniform_losses": "lossy_uniform",
}


def literal(node):
    # also [x] * k, as in np.array([0.9] * 5)
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult):
        return ast.literal_eval(node.left) * ast.literal_eval(node.right)
    return ast.literal_eval(node)


def complex_matrix(call):
    # np.array([[...], ...]) with complex literals
    rows = literal(call.args[0])
    return [[[float(z.real), float(z.imag)] for z in row] for row in rows]


def main():
    tree = ast.parse(open(SRC).read())
    out = {"source": "tests/_simulators/passive/test_measurements.py (reference)", "cases": []}
    for fn in tree.body:
        if not isinstance(fn, ast.FunctionDef) or fn.name not in CASES:
            continue
        case = {"name": fn.name, "kind": CASES[fn.name], "line": fn.lineno}
        for st in ast.walk(fn):
            if isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Name):
                name = st.targets[0].id
                if name == "seed_sequence":
                    case["seed"] = literal(st.value)
                elif name == "U":
                    case["U"] = complex_matrix(st.value)
                elif name == "singular_values":
                    case["singular_values"] = literal(st.value.args[0])
                elif name == "expected_samples":
                    case["expected"] = literal(st.value)
            if isinstance(st, ast.Call) and getattr(st.func, "attr", "") == "NumberState":
                case["input"] = literal(st.args[0])
            if isinstance(st, ast.Call) and getattr(st.func, "attr", "") == "execute":
                for kw in st.keywords:
                    if kw.arg == "shots":
                        case["shots"] = literal(kw.value)
        # every lossy case builds its matrix as U diag(s) U U^T (test_measurements.py:446,529)
        if "singular_values" in case:
            case["lossy_matrix"] = "U @ diag(singular_values) @ U @ U.T"
        out["cases"].append(case)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(out['cases'])} cases to {OUT}")


if __name__ == "__main__":
    main()
