"""Convert the reference's hex8 regression cases into small JSON fixtures.

Reads DATA files only (meshes, boundary conditions, load tables, golden result tables and the
Newton residual logs) from the reference tree and writes them as plain JSON next to this
script.  No reference code is imported or run.  Run once, in a container that has the
reference mounted:

    python tests/golden/make_checker_fixtures.py /root/reference

Cases (paths relative to the reference root):
  Checker/Cases/NonLinear_FEA_3D_Hexa8_StVenant           -> svk_hex8_case.json
      GiD 12 mesh (96 hex8, 195 nodes), St.Venant-Kirchhoff E=100e9 nu=0.3, nonlinear,
      full integration, 50 steps; Marc displacements of node 11 (marc_disp_{x,y,z}.dat,
      checked by the reference at 1e-5 absolute, Run_Checker.bat) and the Newton residual
      history printed by CEOS (Log_CEOS.txt).
  Checker/Cases/Linear_FEA_3D_Hexa8_GeneralizedHookesLaw  -> ghl_hex8_case.json
      HyperMesh .cdb mesh (250 hex8), generalized Hooke's law E=100e9 nu=0.3, linear,
      1 step; CEOS displacements of node 342 and Gauss-point stresses of element 169, GP 5
      (ref_disp_*.dat, ref_stress_*.dat).
  Checker/Cases/NonLinear_FEA_3D_Hexa8_NeoHookean_Incompressible -> nhq1p0_hex8_case.json
      the StVenant case's mesh with its own supports and loads, Neo-Hookean C10 = 5e10,
      K = 5e13 under the mean-dilatation (Q1P0) element technology, 50 steps; Marc
      displacements of node 11 (checked by the reference at 1e-3) and the residual log.
The raw data files of the three cases (inputs, Marc or CEOS results, logs) are copied into
checker/<case>/ for the tests of the package's case reader and runner.
"""
import json
import os
import re
import shutil
import sys

# the input and result files of the three hex8 cases, copied as they are into checker/<case>/ so
# that the package's case reader and runner (ceos_amd.caseio, ceos_amd.analysis) are tested on
# the reference's own files
_GID_FILES = ["Settings.dat", "Mesh.dat", "Time_Discretization.dat", "Table_Disp_Z.tab", "Table_Load_Y.tab",
              "Post_Processing.post", "marc_disp_x.dat", "marc_disp_y.dat", "marc_disp_z.dat", "Log_CEOS.txt",
              "NumOfIterRef.dat"]
RAW_CASES = {
    "NonLinear_FEA_3D_Hexa8_StVenant": _GID_FILES,
    "NonLinear_FEA_3D_Hexa8_NeoHookean_Incompressible": _GID_FILES,
    "Linear_FEA_3D_Hexa8_GeneralizedHookesLaw": [
        "Settings.dat", "Mesh.cdb", "Time_Discretization.dat", "Table_Disp_Z.tab", "Post_Processing.post",
        "ref_disp_x.dat", "ref_disp_y.dat", "ref_disp_z.dat", "ref_stress_11.dat", "ref_stress_33.dat",
        "ref_stress_12.dat", "Log_CEOS.txt"],
}


def _table(path):
    rows = []
    with open(path) as f:
        next(f)
        for line in f:
            parts = line.replace("d", "e").replace("D", "E").split()
            if len(parts) >= 2:
                rows.append([float(parts[0]), float(parts[1])])
    return rows


def _time_discretization(path):
    with open(path) as f:
        toks = [ln.split() for ln in f if ln.strip()]
    n = int(toks[1][0])
    cases = []
    for t in toks[3:3 + n]:
        cases.append({"t0": float(t[1]), "t1": float(t[2]), "steps": int(t[3])})
    return cases


def _newton_log(path):
    """Per step: list of max-abs residual norms printed as 'IT: k NORM: x'."""
    steps = []
    cur = None
    with open(path) as f:
        for line in f:
            if re.match(r"\s*Step:\s*\d+", line):
                cur = []
                steps.append(cur)
            m = re.match(r"\s*IT:\s*(\d+)\s+NORM:\s*([-+0-9.Ee]+)", line)
            if m and cur is not None:
                cur.append(float(m.group(2)))
    return steps


def _gid_case(d):
    """Mesh, material ids, supports, nodal loads / displacements and their tables of a GiD 12
    Mesh.dat (the reader's format, FEM/Input/ModReadInputFile.f90)."""
    lines = open(os.path.join(d, "Mesh.dat")).read().splitlines()
    i = 0
    coords, conn, mat = [], [], []
    fixed, loads, disps = [], [], []
    while i < len(lines):
        s = lines[i].strip().upper()
        if s == "COORDINATES":
            i += 1
            while lines[i].strip().upper() != "END COORDINATES":
                p = lines[i].split()
                coords.append([float(v) for v in p[1:4]])
                i += 1
        elif s == "MATERIAL ID":
            n = int(lines[i + 1].split()[0])
            for k in range(n):
                mat.append(int(lines[i + 2 + k].split()[1]))
            i += 2 + n
            continue
        elif s == "CONNECTIVITY":
            i += 1
            while lines[i].strip().upper() != "END CONNECTIVITY":
                p = [int(v) for v in lines[i].split()]
                conn.append(p[3:3 + p[2]])
                i += 1
        elif s == "FIXED SUPPORTS":
            n = int(lines[i + 1].split()[0])
            for k in range(n):
                p = [int(v) for v in lines[i + 2 + k].split()]
                fixed.append(p)
            i += 2 + n
            continue
        elif s == "NODAL LOAD":
            n = int(lines[i + 1].split()[0])
            for k in range(n):
                loads.append(lines[i + 2 + k].split())
            i += 2 + n
            continue
        elif s == "NODAL DISPLACEMENT":
            n = int(lines[i + 1].split()[0])
            for k in range(n):
                disps.append(lines[i + 2 + k].split())
            i += 2 + n
            continue
        i += 1

    tables = {}

    def bc(entries):
        out = []
        for p in entries:
            node = int(p[0])
            for k in range(3):
                flag, tab = int(p[1 + 2 * k]), p[2 + 2 * k]
                if flag:
                    if tab not in tables:
                        tables[tab] = _table(os.path.join(d, tab + ".tab"))
                    out.append([node, k + 1, tab])
        return out

    return {
        "coords": coords,
        "conn": conn,
        "mat": mat,
        "fixed": [[p[0]] + p[1:4] for p in fixed],
        "nodal_load": bc(loads),
        "nodal_disp": bc(disps),
        "tables": tables,
    }


def _marc(d):
    return {c: _table(os.path.join(d, "marc_disp_%s.dat" % c)) for c in "xyz"}


def svk_case(root):
    d = os.path.join(root, "Checker/Cases/NonLinear_FEA_3D_Hexa8_StVenant")
    g = _gid_case(d)
    return {
        "source": "Checker/Cases/NonLinear_FEA_3D_Hexa8_StVenant",
        "material": {"model": 5, "E": 100.0e9, "nu": 0.3},
        "nl_analysis": True,
        "newton": {"tol": 1e-3, "maxiter": 10},
        "time": _time_discretization(os.path.join(d, "Time_Discretization.dat")),
        "coords": g["coords"],
        "conn": g["conn"],
        "mat": g["mat"],
        "fixed": g["fixed"],
        "nodal_load": g["nodal_load"],
        "nodal_disp": g["nodal_disp"],
        "tables": g["tables"],
        "probe_node": 11,
        "marc_disp": _marc(d),
        "marc_tol": 1.0e-5,
        "log_norms": _newton_log(os.path.join(d, "Log_CEOS.txt")),
    }


def nhq1p0_case(root):
    """Settings.dat: Neo_Hookean C10 = 5e10, bulk modulus 5e13, Element Technology = Mean
    Dilatation (the reference selects model 4, NeoHookeanQ1P0, for that pair,
    ModConstitutiveModelLibrary.f90:504), nonlinear, Newton tol 1e-3, 10 iterations;
    Run_Checker.bat compares the node-11 probes with Marc at 1e-3."""
    d = os.path.join(root, "Checker/Cases/NonLinear_FEA_3D_Hexa8_NeoHookean_Incompressible")
    g = _gid_case(d)
    out = {
        "source": "Checker/Cases/NonLinear_FEA_3D_Hexa8_NeoHookean_Incompressible",
        "material": {"model": 4, "C10": 5.0e10, "K": 5.0e13},
        "element_technology": "mean dilatation",
        "nl_analysis": True,
        "newton": {"tol": 1e-3, "maxiter": 10},
        "time": _time_discretization(os.path.join(d, "Time_Discretization.dat")),
        "probe_node": 11,
        "marc_disp": _marc(d),
        "marc_tol": 1.0e-3,
        "log_norms": _newton_log(os.path.join(d, "Log_CEOS.txt")),
    }
    out.update(g)
    return out


def ghl_case(root):
    d = os.path.join(root, "Checker/Cases/Linear_FEA_3D_Hexa8_GeneralizedHookesLaw")
    lines = open(os.path.join(d, "Mesh.cdb")).read().splitlines()
    coords, conn = {}, {}
    collector = None
    fixed, presc = {}, {}
    i = 0
    while i < len(lines):
        ln = lines[i]
        up = ln.strip().upper()
        if up.startswith("NBLOCK"):
            i += 2
            while not lines[i].strip().upper().startswith("N,"):
                p = lines[i].split()
                coords[int(p[0])] = [float(v) for v in p[3:6]]
                i += 1
            continue
        if up.startswith("EBLOCK"):
            i += 2
            while lines[i].split()[0] != "-1":
                p = [int(v) for v in lines[i].split()]
                conn[p[10]] = p[11:19]
                i += 1
            continue
        m = re.match(r"!!\s+(\d+)\s+\"(.*)\"", ln.strip())
        if m and "LOADCOL" in lines[i - 1].upper():
            collector = m.group(2)
        if up.startswith("D,"):
            p = [v.strip() for v in ln.split(",")]
            node, dof = int(p[1]), "XYZ".index(p[2].upper()[-1]) + 1
            if collector == "Fixed Supports":
                fixed.setdefault(node, [0, 0, 0])[dof - 1] = 1
            else:
                presc.setdefault(node, {})[dof] = collector
        i += 1
    nn = max(coords)
    ne = max(conn)
    tables = {}
    nodal_disp = []
    for node in sorted(presc):
        for dof, tab in sorted(presc[node].items()):
            if tab not in tables:
                tables[tab] = _table(os.path.join(d, tab + ".tab"))
            nodal_disp.append([node, dof, tab])
    ref = {}
    for name in ["disp_x", "disp_y", "disp_z", "stress_11", "stress_33", "stress_12"]:
        ref[name] = _table(os.path.join(d, "ref_%s.dat" % name))
    return {
        "source": "Checker/Cases/Linear_FEA_3D_Hexa8_GeneralizedHookesLaw",
        "material": {"model": 1, "E": 100.0e9, "nu": 0.3},
        "nl_analysis": False,
        "newton": {"tol": 1e-3, "maxiter": 10},
        "time": _time_discretization(os.path.join(d, "Time_Discretization.dat")),
        "coords": [coords[k] for k in range(1, nn + 1)],
        "conn": [conn[k] for k in range(1, ne + 1)],
        "mat": [1] * ne,
        "fixed": [[n] + fixed[n] for n in sorted(fixed)],
        "nodal_load": [],
        "nodal_disp": nodal_disp,
        "tables": tables,
        "probe_node": 342,
        "probe_elem": 169,
        "probe_gp": 5,
        "ref": ref,
    }


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    for name, fn in [("svk_hex8_case.json", svk_case), ("ghl_hex8_case.json", ghl_case),
                     ("nhq1p0_hex8_case.json", nhq1p0_case)]:
        data = fn(root)
        with open(os.path.join(here, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print("wrote", name, "nodes", len(data["coords"]), "elems", len(data["conn"]))
    for case, files in RAW_CASES.items():
        src = os.path.join(root, "Checker/Cases", case)
        dst = os.path.join(here, "checker", case)
        os.makedirs(dst, exist_ok=True)
        for fn in files:
            shutil.copyfile(os.path.join(src, fn), os.path.join(dst, fn))
        print("copied", case)


if __name__ == "__main__":
    main()
