"""Readers and writers of the reference's analysis files, so that its hex8 cases run from their
own input files through the library.

  Settings.dat              FEM/Input/ModReadInputFile.f90 (sections ANALYSIS SETTINGS, NONLINEAR
                            SOLVER, MATERIAL, MESH AND BOUNDARY CONDITIONS): case insensitive,
                            blanks ignored, '!' starts a comment, 'name = value' options
  Mesh.dat (Preprocessor = Gid 12)
                            ModReadInputFile.f90:1396-1714: COORDINATES, MATERIAL ID,
                            CONNECTIVITY (hex8), FIXED SUPPORTS, Nodal Load, Nodal Displacement
                            (per node three flag / table pairs)
  Mesh.cdb (Preprocessor = HyperMesh)
                            ReadMeshHyperMesh / ReadBoundaryConditionsHyperMesh,
                            ModReadInputFile.f90:909-1395: the ANSYS deck HyperMesh writes
                            (NUMOFF, NBLOCK, ET 185 = hex8, EBLOCK), then one load collector per
                            table ('Fixed Supports', or the table's name for its D and F lines)
  <table>.tab               ModLoadHistoryData.f90:122-223: a header line, then 'time value' rows,
                            linearly interpolated
  Time_Discretization.dat   load cases: initial time, final time, number of steps
  Post_Processing.post      the probes (ceos_amd.probes, FEM/PostProcessing/ModProbe.f90)
  FEMAnalysis.result        WriteFEMResults (FEM/ModFEMAnalysis.f90:1094-1129), one record per
                            converged (sub)step after a 'Total Number of DOF' line; replayed for
                            the probes as ModExportResultFile.f90:254-440 does
Material names map to the library's model ids as ModConstitutiveModelLibrary.f90:492-556 pairs
them with the element technology; option names follow ReadMaterialParameters_* of each model.
"""
import os

import numpy as np

from . import probes as _probes

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
    multiscale = _key(an.get("multiscaleanalysis", "false"))
    if multiscale not in ("true", "false"):
        raise CaseError("Multiscale Analysis not identified")
    multiscale = multiscale == "true"
    if multiscale and _key(an.get("multiscalemodel", "")) != "periodic":
        raise CaseError("multiscale model %r is not run: only Periodic" % an.get("multiscalemodel", ""))
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
    pre = _key(mb.get("preprocessor", "gid12"))
    if pre not in ("gid12", "hypermesh"):
        raise CaseError("only GiD 12 and HyperMesh meshes are read")
    macro = None
    if multiscale:
        # MACROSCOPIC DEFORMATION GRADIENT (ReadMacroscopicDeformationGradient,
        # ModReadInputFile.f90:437-560): F11 ... F33, each 'Zero', 'One' or a table name
        mf = _options(sec.get("macroscopicdeformationgradient", []))
        macro = {}
        for i in range(3):
            for j in range(3):
                k = "f%d%d" % (i + 1, j + 1)
                if k not in mf:
                    raise CaseError("Multiscale Settings :: Option not found [F%d%d]" % (i + 1, j + 1))
                v = mf[k].strip()
                macro[(i, j)] = 0.0 if _key(v) == "zero" else (1.0 if _key(v) == "one" else v)
    return {
        "multiscale": multiscale,
        "macro_F": macro,
        "nl_analysis": nl,
        "element_technology": tech,
        "max_cut_backs": int(_num(an.get("maximumcutbacks", "10"))),
        "restart": restart == "true",
        "newton_tol": _num(ns.get("tol", "1e-3")),
        "newton_maxiter": int(_num(ns.get("maxiter", "10"))),
        "materials": mats,
        "preprocessor": pre,
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
            "mat": np.array(mat, dtype=np.int32), "fixed": fixed, "nodal_load": loads, "nodal_disp": disps,
            "boundary": []}


def read_hypermesh_mesh(path):
    """The mesh and boundary conditions of a HyperMesh ANSYS deck, in the dictionary
    read_gid_mesh returns.  As ReadMeshHyperMesh (ModReadInputFile.f90:909-1080): the NUMOFF
    lines give the node and element counts; NBLOCK is followed by a format line and one line per
    node, numbered in file order, whose 4th to 6th blank-separated fields are x, y, z; ET type 185
    is the hex8; EBLOCK is followed by a format line and one line per element up to '-1', with
    the material id in field 1, the node count in field 9, the element id in field 11 and the
    nodes from field 12.  From the '!! Loadstep Data' line on, as
    ReadBoundaryConditionsHyperMesh (:1182-1395): every '!!HMNAME LOADCOL' line is followed by the
    collector's name in double quotes, and every later line of four comma-separated fields
    belongs to it; 'Fixed Supports' lists the fixed DOFs (D, node, UX|UY|UZ), collectors named
    'boundary...' (RVE node sets) and 'reaction...' are skipped, and any other collector names the
    table (<name>.tab) of its prescribed displacements (D lines) and nodal loads (F, node,
    FX|FY|FZ)."""
    lines = open(path, encoding="latin-1").read().splitlines()
    nn = ne = None
    node_lines, elems, etype = None, {}, None
    i = 0
    while i < len(lines):
        line = lines[i]
        f = line.split(",")
        if _key(line) == "!!loadstepdata":
            i += 1
            break
        if len(f) >= 3 and _key(f[0]) == "numoff":
            if _key(f[1]) == "node":
                nn = int(f[2])
            elif _key(f[1]) == "elem":
                ne = int(f[2])
        elif len(f) >= 3 and _key(f[0]) == "nblock":
            if nn is None:
                raise CaseError("NBLOCK before NUMOFF,NODE in %s" % path)
            node_lines = lines[i + 2:i + 2 + nn]
            i += 2 + nn
            continue
        elif len(f) >= 3 and _key(f[0]) == "et":
            etype = int(f[2])
            if etype != 185:
                raise CaseError("ANSYS element type %d is not supported: only 185 (hex8)" % etype)
        elif len(f) >= 3 and _key(f[0]) == "eblock":
            i += 2
            while i < len(lines) and lines[i].split() and lines[i].split()[0] != "-1":
                p = [int(v) for v in lines[i].split()]
                if p[8] != 8:
                    raise CaseError("element %d has %d nodes: only hex8 is supported" % (p[10], p[8]))
                elems[p[10]] = (p[0], p[11:19])
                i += 1
        i += 1
    if nn is None or ne is None or node_lines is None or etype is None:
        raise CaseError("no NUMOFF, NBLOCK or ET line in %s" % path)
    if len(node_lines) != nn or sorted(elems) != list(range(1, ne + 1)):
        raise CaseError("%s: the node or element blocks do not match NUMOFF" % path)
    coords = np.array([[_num(v) for v in ln.split()[3:6]] for ln in node_lines])
    conn = np.array([elems[k][1] for k in range(1, ne + 1)], dtype=np.int32)
    mat = np.array([elems[k][0] for k in range(1, ne + 1)], dtype=np.int32)
    fixed, loads, disps = {}, [], []
    boundary = []
    name = None
    dof_of = {"ux": 1, "uy": 2, "uz": 3, "fx": 1, "fy": 2, "fz": 3}
    while i < len(lines):
        line = lines[i]
        if _key(line) == "!!hmnameloadcol" and i + 1 < len(lines):
            q = lines[i + 1].split('"')
            name = q[1] if len(q) > 1 else ""
            i += 2
            continue
        f = [v.strip() for v in line.split(",")]
        if len(f) == 4 and name is not None:
            k = _key(name)
            first = name.split()[0].lower() if name.split() else ""
            if first == "boundary":
                # node sets of an RVE's boundary (BC%BoundaryNodes, ModReadInputFile.f90:1290-1300)
                if not boundary or boundary[-1][0] != name:
                    boundary.append((name, []))
                boundary[-1][1].append(int(f[1]))
            elif first == "reaction":
                pass
            elif k == "fixedsupports":
                d = dof_of.get(f[2].lower())
                if d is None or d > 3 or f[2].lower()[0] != "u":
                    raise CaseError("DOF %r not identified in %s" % (f[2], path))
                fixed.setdefault(int(f[1]), [0, 0, 0])[d - 1] = 1
            elif _key(f[0]) == "d":
                if f[2].lower() not in ("ux", "uy", "uz"):
                    raise CaseError("DOF %r not identified in %s" % (f[2], path))
                disps.append([int(f[1]), dof_of[f[2].lower()], name])
            elif _key(f[0]) == "f":
                if f[2].lower() not in ("fx", "fy", "fz"):
                    raise CaseError("DOF %r not identified in %s" % (f[2], path))
                loads.append([int(f[1]), dof_of[f[2].lower()], name])
        i += 1
    return {"coords": coords, "conn": conn, "mat": mat, "fixed": [[n] + fixed[n] for n in sorted(fixed)],
            "nodal_load": loads, "nodal_disp": disps, "boundary": boundary}


def table_path(directory, name):
    """File of a load table (ReadValueDiscretization, ModLoadHistoryData.f90:140-148): the name
    itself when it has an extension, else name + '.tab'."""
    return os.path.join(directory, name if "." in name else name + ".tab")


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


def interp(table, t):
    ts = [p[0] for p in table]
    vs = [p[1] for p in table]
    return float(np.interp(t, ts, vs))


class Case:
    """A hex8 case read from a directory holding Settings.dat and the files it names."""

    def __init__(self, directory, settings="Settings.dat"):
        self.dir = directory
        self.settings = read_settings(os.path.join(directory, settings))
        mesh_path = os.path.join(directory, self.settings["mesh_file"])
        if self.settings["preprocessor"] == "hypermesh":
            self.mesh = read_hypermesh_mesh(mesh_path)
        else:
            self.mesh = read_gid_mesh(mesh_path)
        self.load_cases = read_time_discretization(os.path.join(directory, self.settings["time_file"]))
        self.probes = _probes.read_probes(os.path.join(directory, "Post_Processing.post"), directory)
        self.tables = {}
        names = [tab for _, _, tab in self.mesh["nodal_load"] + self.mesh["nodal_disp"]]
        if self.settings["multiscale"]:
            names += [v for v in self.settings["macro_F"].values() if isinstance(v, str)]
        for tab in names:
            if tab not in self.tables:
                self.tables[tab] = read_table(table_path(directory, tab))
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

    def macro_F(self, t):
        """The macroscopic deformation gradient of a multiscale case at time t (each component a
        constant or its table, linearly interpolated)."""
        F = np.zeros((3, 3))
        for (i, j), v in self.settings["macro_F"].items():
            F[i, j] = interp(self.tables[v], t) if isinstance(v, str) else v
        return F

    def boundary_nodes(self):
        """0-based nodes of the mesh's first boundary set (BC%BoundaryNodes(1)), or None."""
        b = self.mesh.get("boundary") or []
        return np.asarray(b[0][1], dtype=np.int64) - 1 if b else None


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


def _homogenized(mesh):
    """HomogenizeStress / HomogenizeDeformationGradient of the replayed RVE (the volume averages of
    P and F over its Gauss points): (P 9 column-major, F 3x3)."""
    def get():
        P, F = mesh.homogenize()
        return P, np.asarray(F).reshape(3, 3).T
    return get


def post_process(case, results_path, out_dir, mesh_factory=None):
    """PostProcessingResults (ModExportResultFile.f90:254-440): the probe files start with their
    header; every record of FEMAnalysis.result, in file order, sets the time and U and, when a
    probe reads the Gauss points or the internal force (or, in a multiscale case, the homogenized
    stress and deformation gradient of the RVE), is replayed through a fresh mesh (the
    mesh moves to X + U in a finite-strain analysis, the constitutive model is updated at the
    record's time, and the state is committed after the record); the records whose FLAG END
    STEP is 1 write one row of every active probe and of NumberOfIterationsToConverge.dat.
    mesh_factory(case) makes the replay mesh (default: the library's, as the analysis runs on).
    Returns {probe file: rows as written, (time, values) or message}."""
    plist = case.probes
    for p in plist:
        p.active = True
        p.rows = []
        p.initialize(out_dir)
    ndof, recs = read_results(results_path)
    if ndof != case.ndof:
        raise CaseError("FEMAnalysis.result holds %d DOFs, the case %d" % (ndof, case.ndof))
    mesh = None
    multiscale = case.settings.get("multiscale", False)
    if any(p.needs_mesh or (multiscale and p.location == "microstructure") for p in plist):
        if mesh_factory is None:
            from .analysis import make_mesh as mesh_factory
        mesh = mesh_factory(case)
    nl = case.settings["nl_analysis"]
    nn, ne = case.ndof // 3, len(case.mesh["conn"])
    try:
        with open(os.path.join(out_dir, "NumberOfIterationsToConverge.dat"), "w") as itf:
            itf.write("   Time                    Number Of Iterations To Converge\n")
            for r in recs:
                if mesh is not None:
                    if nl:
                        mesh.update_mesh(r["U"])
                    mesh.solve_constitutive(r["time"], r["U"])
                if r["flag_end_step"] == 1:
                    rec = _probes.Record(r["time"], r["U"], nn, ne, mesh, case.settings["materials"],
                                         case.mesh["mat"], homogenized=_homogenized(mesh) if multiscale else None)
                    for p in plist:
                        p.write(out_dir, rec)
                    itf.write(" %24.15E %11d\n" % (r["time"], r["iterations"]))
                if mesh is not None:
                    mesh.switch_converged_state()
    finally:
        if mesh is not None:
            mesh.close()
    return {p.file: p.rows for p in plist}
