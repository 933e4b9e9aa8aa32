"""Extract the reference's blessed regression outputs into a JSON fixture.

Reads /root/reference/tests/<name>.prm and <name>.output (the reference's own
test data: Newton iteration counts and point displacements per load step,
written by Solid::run / print_solution, include/mf_elasticity.h:1187-1222,
2615-2690) and writes tests/golden/reference_outputs.json.  Only data is
kept: the settings of the prm files that define the problem, and the
numbers of the output files.  Run in the build container; the JSON is the
committed fixture.  The parameter files and blessed outputs themselves (input
and expected-output data of those tests) are copied to tests/golden/prm/ so
the parameter-file driver can be run on them unchanged.
"""
import json
import os
import re
import shutil

REF = "/root/reference/tests"
NAMES = ["cook_mf_form_1", "cook_mat_form_1", "cook_mf_gmg_form_1", "cook_mf_form_1_tensor4_ns",
         "cook_mf", "cook", "cook_mf_gmg", "cook_3d_quad_mf", "cook_3d_lin_mf"]
KEYS = ["Polynomial degree", "Quadrature order", "Elements per edge", "Global refinement", "Grid scale",
        "Dimension", "Type", "Dirichlet IDs and expressions", "Dirichlet IDs and component mask",
        "Neumann IDs and expressions", "Output points", "Residual", "Preconditioner type",
        "Solver type", "MF caching", "Poisson's ratio", "Shear modulus", "Formulation",
        "Max iterations Newton-Raphson", "Tolerance displacement", "Tolerance force", "End time",
        "Time step size"]


def parse_prm(path):
    out = {}
    for line in open(path):
        line = line.split("#")[0].strip()
        m = re.match(r"set\s+(.+?)\s*=\s*(.*)$", line)
        if m and m.group(1) in KEYS:
            out[m.group(1)] = m.group(2).strip()
    return out


def parse_output(path):
    txt = open(path).read()
    res = {"steps": []}
    m = re.search(r"Reference volume:\s*(\S+)", txt)
    res["reference_volume"] = float(m.group(1))
    res["n_cells"] = int(re.search(r"Number of active cells:\s*(\d+)", txt).group(1))
    res["n_dofs"] = int(re.search(r"Number of degrees of freedom:\s*(\d+)", txt).group(1))
    for block in re.split(r"\nTimestep ", txt)[1:]:
        step = {}
        m = re.match(r"(\d+) @ (\S+)s", block)
        step["step"] = int(m.group(1))
        step["time"] = float(m.group(2))
        m = re.search(r"Converged in (\d+) Newton iterations", block)
        step["newton"] = int(m.group(1)) if m else None
        m = re.search(r"Solution @ ([^\n]+)\n\s*displacement:\s*([^\n]+)", block)
        if m:
            step["point"] = [float(x) for x in m.group(1).split()]
            step["displacement"] = [float(x) for x in m.group(2).split()]
        res["steps"].append(step)
    return res


def main():
    data = {}
    for n in NAMES:
        prm, out = os.path.join(REF, n + ".prm"), os.path.join(REF, n + ".output")
        if os.path.exists(prm) and os.path.exists(out):
            data[n] = {"prm": parse_prm(prm), "output": parse_output(out)}
    here = os.path.dirname(os.path.abspath(__file__))
    os.makedirs(os.path.join(here, "prm"), exist_ok=True)
    for n in NAMES:
        for ext in (".prm", ".output"):
            src = os.path.join(REF, n + ext)
            if os.path.exists(src):
                shutil.copyfile(src, os.path.join(here, "prm", n + ext))
    with open(os.path.join(here, "reference_outputs.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote", len(data), "cases")


if __name__ == "__main__":
    main()
