"""Readers and writers of the reference's analysis files, so that its hex8 cases run from their
own input files through the library.

  Settings.dat              FEM/Input/ModReadInputFile.f90 (sections ANALYSIS SETTINGS, NONLINEAR
                            SOLVER, MATERIAL, MESH AND BOUNDARY CONDITIONS): case insensitive,
                            blanks ignored, '!' starts a comment, 'name = value' options
  Mesh.dat (Preprocessor = Gid 12)
                            ModReadInputFile.f90:1396-1714: COORDINATES, MATERIAL ID,
                            CONNECTIVITY (hex8), FIXED SUPPORTS, Nodal Load, Nodal Displacement
                            (per node three flag / table pairs)
  <table>.tab               ModLoadHistoryData.f90:122-223: a header line, then 'time value' rows,
                            linearly interpolated
  Time_Discretization.dat   load cases: initial time, final time, number of steps
  Post_Processing.post      node probes of the displacements (FEM/PostProcessing/ModProbe.f90)
  FEMAnalysis.result        WriteFEMResults (FEM/ModFEMAnalysis.f90:1094-1129), one record per
                            converged (sub)step after a 'Total Number of DOF' line; read back for
                            the probes as ModExportResultFile.f90:340-420 does
  probe files               ModProbe.f90:170-206: header ' Time                    Value', then
                            rows in E23.15
Material names map to the library's model ids as ModConstitutiveModelLibrary.f90:492-556 pairs
them with the element technology; option names follow ReadMaterialParameters_* of each model.
"""
import os
import re

import numpy as np

# (name, element technology) -> model id, option names in the order of the library's props
MATERIALS = {
    ("generalized_hookes_law", 1): (1, ["youngmodulus", "poisson"]),
    ("j2_plasticity", 1): (2, ["youngmodulus", "poisson", "yieldstress", "hardeningmodulus"]),
    ("neo_hookean", 1): (3, ["c10", "bulkmodulus"]),
    ("neo_hookean", 2): (4, ["c10", "bulkmodulus"]),
    ("st_venant_kirchhoff", 1): (5, ["youngmodulus", "poisson"]),
    ("compressible_neo_hookean", 1): (7, ["mu", "lambda"]),
    ("matrix_viscoelastic", 1): (12, ["k_inf", "mu_inf", "lambda_inf", "k_e", "mu_e", "ni_v"]),
}
ELEMENT_TECHNOLOGY = {"fullintegration": 1, "meandilatation": 2}


class CaseError(ValueError):
    pass


def _key(s):
    """Case- and blank-insensitive form of an option name or value."""
    return "".join(s.split()).lower()


def _num(s):
    return float(s.strip().replace("d", "e").replace("D", "E"))


def _lines(path):
    """Lines without comments and blank lines."""
    out = []
    with open(path, encoding="latin-1") as f:
        for line in f:
            line = line.split("!", 1)[0].strip()
            if line:
                out.append(line)
    return out


def _sections(lines):
    """{section key: [lines]} for 'NAME ... END NAME' blocks."""
    out, cur, name = {}, None, None
    for line in lines:
        k = _key(line)
        if cur is None:
            name, cur = k, []
        elif k == "end" + name:
            out[name] = cur
            cur = None
        else:
            cur.append(line)
    return out


def _options(lines):
    opts = {}
    for line in lines:
        if "=" in line:
            a, b = line.split("=", 1)
            opts[_key(a)] = b.strip()
    return opts


def read_settings(path):
    """Analysis settings, Newton parameters, materials and file names of a Settings.dat."""
    sec = _sections(_lines(path))
    an = _options(sec.get("analysissettings", []))
    if _key(an.get("multiscaleanalysis", "false")) != "false":
        raise CaseError("multiscale analyses are not read")
    if _key(an.get("hypothesisofanalysis", "3d")) != "3d":
        raise CaseError("only 3D analyses are supported")
    tech = ELEMENT_TECHNOLOGY.get(_key(an.get("elementtechnology", "fullintegration")))
    if tech is None:
        raise CaseError("element technology %r is not supported" % an.get("elementtechnology"))
    nl = _key(an.get("nonlinearanalysis", "true")) == "true"
    # Restart (ModReadInputFile.f90:287-301): True resumes from the FEMAnalysis.result of the case
    restart = _key(an.get("restart", "false"))
    if restart not in ("true", "false"):
        raise CaseError("Restart option not identified or not implemented")
    ns = _options(sec.get("nonlinearsolver", []))
    mats = {}
    lines = sec.get("material", [])
    i = 0
    while i < len(lines):
        a = lines[i]
        if _key(a).startswith("materialid="):
            mid = int(_num(a.split("=", 1)[1]))
            name = _key(lines[i + 1])
            if (name, tech) not in MATERIALS:
                raise CaseError("material %r with element technology %d is not supported" % (lines[i + 1], tech))
            model, names = MATERIALS[(name, tech)]
            opts = {}
            j = i + 2
            while j < len(lines) and not _key(lines[j]).startswith("materialid="):
                if "=" in lines[j]:
                    k, v = lines[j].split("=", 1)
                    opts[_key(k)] = _num(v)
                j += 1
            missing = [n for n in names if n not in opts]
            if missing:
                raise CaseError("material %d: options %s not found" % (mid, missing))
            mats[mid] = (model, [opts[n] for n in names])
            i = j
            continue
        i += 1
    mb = _options(sec.get("meshandboundaryconditions", []))
    if _key(mb.get("preprocessor", "gid12")) != "gid12":
        raise CaseError("only GiD 12 meshes are read")
    return {
        "nl_analysis": nl,
        "element_technology": tech,
        "max_cut_backs": int(_num(an.get("maximumcutbacks", "10"))),
        "restart": restart == "true",
        "newton_tol": _num(ns.get("tol", "1e-3")),
        "newton_maxiter": int(_num(ns.get("maxiter", "10"))),
        "materials": mats,
        "mesh_file": mb.get("meshfile", "Mesh.dat"),
        "time_file": mb.get("timediscretizationfile", "Time_Discretization.dat"),
    }


def read_gid_mesh(path):
    """Coordinates, hex8 connectivity (1-based), material ids and the boundary-condition lists
    of a GiD 12 Mesh.dat: fixed supports [node, fx, fy, fz]; nodal loads and displacements
    [node, dof (1..3), table name] for every flag set."""
    lines = [ln.strip() for ln in open(path, encoding="latin-1").read().splitlines()]
    coords, conn, mat, fixed, loads, disps = [], [], [], [], [], []
    i = 0
    while i < len(lines):
        k = _key(lines[i])
        if k == "coordinates":
            i += 1
            while _key(lines[i]) != "endcoordinates":
                if lines[i]:
                    p = lines[i].split()
                    coords.append([_num(v) for v in p[1:4]])
                i += 1
        elif k == "materialid":
            n = int(lines[i + 1].split()[0])
            mat = [int(lines[i + 2 + j].split()[1]) for j in range(n)]
            i += 2 + n
            continue
        elif k == "connectivity":
            i += 1
            while _key(lines[i]) != "endconnectivity":
                if lines[i]:
                    p = [int(v) for v in lines[i].split()]
                    if p[2] != 8:
                        raise CaseError("element %d has %d nodes: only hex8 is supported" % (p[0], p[2]))
                    conn.append(p[3:11])
                i += 1
        elif k == "fixedsupports":
            n = int(lines[i + 1].split()[0])
            fixed = [[int(v) for v in lines[i + 2 + j].split()[:4]] for j in range(n)]
            i += 2 + n
            continue
        elif k in ("nodalload", "nodaldisplacement"):
            n = int(lines[i + 1].split()[0])
            out = loads if k == "nodalload" else disps
            for j in range(n):
                p = lines[i + 2 + j].split()
                node = int(p[0])
                for d in range(3):
                    if int(p[1 + 2 * d]):
                        out.append([node, d + 1, p[2 + 2 * d]])
            i += 2 + n
            continue
        i += 1
    if not coords or not conn:
        raise CaseError("no coordinates or connectivity in %s" % path)
    if not mat:
        mat = [1] * len(conn)
    return {"coords": np.array(coords), "conn": np.array(conn, dtype=np.int32),
            "mat": np.array(mat, dtype=np.int32), "fixed": fixed, "nodal_load": loads, "nodal_disp": disps}


def read_table(path):
    rows = []
    with open(path, encoding="latin-1") as f:
        next(f)
        for line in f:
            p = line.split()
            if len(p) >= 2:
                rows.append((_num(p[0]), _num(p[1])))
    return rows


def read_time_discretization(path):
    t = [ln.split() for ln in _lines(path)]
    n = int(t[1][0])
    return [{"t0": _num(r[1]), "t1": _num(r[2]), "steps": int(r[3])} for r in t[3:3 + n]]


def read_probes(path):
    """Node displacement probes of a Post_Processing.post: [{file, node, component}]."""
    out = []
    if not os.path.exists(path):
        return out
    blocks = re.split(r"(?im)^\s*end\s+probe\s*$", open(path, encoding="latin-1").read())
    for b in blocks:
        m = re.search(r"(?im)^\s*probe\s*$", b)
        if not m:
            continue
        opts = _options(_lines_from_text(b[m.end():]))
        if _key(opts.get("location", "")) != "node" or _key(opts.get("variablename", "")) != "displacements":
            continue
        out.append({"file": opts["filename"], "node": int(_num(opts["node"])),
                    "component": int(_num(opts["components"]))})
    return out


def _lines_from_text(text):
    return [ln.split("!", 1)[0].strip() for ln in text.splitlines() if ln.split("!", 1)[0].strip()]


def interp(table, t):
    ts = [p[0] for p in table]
    vs = [p[1] for p in table]
    return float(np.interp(t, ts, vs))


class Case:
    """A hex8 case read from a directory holding Settings.dat and the files it names."""

    def __init__(self, directory, settings="Settings.dat"):
        self.dir = directory
        self.settings = read_settings(os.path.join(directory, settings))
        self.mesh = read_gid_mesh(os.path.join(directory, self.settings["mesh_file"]))
        self.load_cases = read_time_discretization(os.path.join(directory, self.settings["time_file"]))
        self.probes = read_probes(os.path.join(directory, "Post_Processing.post"))
        self.tables = {}
        for _, _, tab in self.mesh["nodal_load"] + self.mesh["nodal_disp"]:
            if tab not in self.tables:
                self.tables[tab] = read_table(os.path.join(directory, tab + ".tab"))
        nn = len(self.mesh["coords"])
        self.ndof = 3 * nn
        fx = []
        for row in self.mesh["fixed"]:
            fx += [3 * (row[0] - 1) + k for k in range(3) if row[1 + k]]
        self.fixed_dof = np.array(sorted(set(fx)), dtype=np.int64)
        self.presc = [(3 * (n - 1) + d - 1, tab) for n, d, tab in self.mesh["nodal_disp"]]
        self.loads = [(3 * (n - 1) + d - 1, tab) for n, d, tab in self.mesh["nodal_load"]]
        self.presc_dof = np.array([d for d, _ in self.presc], dtype=np.int64)

    def fext(self, t):
        f = np.zeros(self.ndof)
        for d, tab in self.loads:
            f[d] += interp(self.tables[tab], t)
        return f

    def presc_values(self, t):
        return np.array([interp(self.tables[tab], t) for _, tab in self.presc])


def write_results_header(f, ndof):
    f.write(" Total Number of DOF =  %d\n" % ndof)


def write_results_record(f, U, time, lc, st, cutback, substep, flag_end_step, iterations):
    """One WriteFEMResults record (list-directed output of the reference)."""
    f.write(" TIME = %24.16E\n LOAD CASE = %d\n STEP = %d\n CUT BACK = %d\n SUBSTEP = %d\n"
            " FLAG END STEP = %d\n NUMBER OF ITERATIONS TO CONVERGE = %d\n"
            % (time, lc, st, cutback, substep, flag_end_step, iterations))
    f.write("".join("  %24.16E\n" % v for v in U))


def read_results(path):
    """(ndof, records) of a FEMAnalysis.result; a record is a dict with time, load_case, step,
    cut_back, substep, flag_end_step, iterations and U."""
    with open(path) as f:
        lines = f.read().splitlines()
    ndof = int(lines[0].split("=")[1])
    recs, i = [], 1
    keys = ["time", "load_case", "step", "cut_back", "substep", "flag_end_step", "iterations"]
    while i + 7 <= len(lines):
        vals = [lines[i + k].split("=")[1] for k in range(7)]
        r = {k: (_num(v) if k == "time" else int(_num(v))) for k, v in zip(keys, vals)}
        r["U"] = np.array([_num(v) for v in lines[i + 7:i + 7 + ndof]])
        recs.append(r)
        i += 7 + ndof
    return ndof, recs


def write_probe(path, rows):
    with open(path, "w") as f:
        f.write("  Time                    Value\n")
        for t, v in rows:
            f.write("%23.15E %23.15E \n" % (t, v))


def post_process(case, results_path, out_dir):
    """The probes and NumberOfIterationsToConverge.dat from FEMAnalysis.result: one row per
    record whose FLAG END STEP is 1 (ModExportResultFile.f90:340-420).  Returns {file: rows}."""
    _, recs = read_results(results_path)
    ends = [r for r in recs if r["flag_end_step"] == 1]
    out = {}
    for p in case.probes:
        dof = 3 * (p["node"] - 1) + p["component"] - 1
        rows = [(r["time"], r["U"][dof]) for r in ends]
        write_probe(os.path.join(out_dir, p["file"]), rows)
        out[p["file"]] = rows
    with open(os.path.join(out_dir, "NumberOfIterationsToConverge.dat"), "w") as f:
        f.write("   Time                    Number Of Iterations To Converge\n")
        for r in ends:
            f.write(" %24.15E %11d\n" % (r["time"], r["iterations"]))
    return out
