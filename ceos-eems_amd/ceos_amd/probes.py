"""Probes of the reference's post-processing: the files that Post_Processing.post asks for, one
row per finished step of FEMAnalysis.result.

  input blocks          PROBE ... END PROBE with Location, File Name, Variable Name, Node,
                        Components, Element, Gauss Point, HyperMesh File, Load Collector
                        (FEM/PostProcessing/ModExportResultFile.f90:140-240)
  variable names        ParseVariableName, FEM/PostProcessing/ModProbe.f90:120-141: Cauchy
                        Stress, Deformation Gradient, Displacements, Logarithmic Strain,
                        Temperature, First Piola Stress; any other name is a user-defined
                        result of the material (GetResult of each model)
  file layout           InitializeFile / WriteOnFile_Char / WriteOnFile_Value,
                        ModProbe.f90:148-208: header ' Time                    Value', then the
                        time and the values in E23.15 separated by one blank, or a message
  Node                  WriteProbeResult_Node, ModProbe.f90:263-335
  Gauss Point           WriteProbeResult_GaussPoint, ModProbe.f90:387-563
  Nodal Force           NodalForceProbeConstructor / WriteProbeResult_NodalForce,
                        ModProbe.f90:613-762
  Micro Structure       WriteProbeResult_MicroStructure, ModProbe.f90:769-908

A probe that meets an error writes the reference's message once and stops writing (its Active
flag).  Gauss-point and nodal-force values come from the library: the post-processing replays
the records through it (caseio.post_process); the values here are host arithmetic on what it
returns for one point, as the reference's probe routines are.
"""
import os
import re

import numpy as np

HEADER = "  Time                    Value\n"

# ParseVariableName (ModProbe.f90:120-141), names compared without case and blanks
CAUCHY_STRESS, DEFORMATION_GRADIENT, DISPLACEMENTS = "cauchystress", "deformationgradient", "displacements"
LOGARITHMIC_STRAIN, TEMPERATURE, FIRST_PIOLA_STRESS = "logarithmicstrain", "temperature", "firstpiolastress"
USER_DEFINED = "userdefined"
_KNOWN = (CAUCHY_STRESS, DEFORMATION_GRADIENT, DISPLACEMENTS, LOGARITHMIC_STRAIN, TEMPERATURE, FIRST_PIOLA_STRESS)

LOCATIONS = ("node", "gausspoint", "nodalforce", "microstructure")
_OPTIONS = {"location": "location", "filename": "file", "variablename": "variable", "node": "node",
            "components": "components", "element": "element", "gausspoint": "gauss_point",
            "hypermeshfile": "hypermesh_file", "loadcollector": "load_collector"}


class ProbeError(ValueError):
    pass


def _key(s):
    return "".join(str(s).split()).lower()


def variable_id(name):
    k = _key(name)
    return k if k in _KNOWN else USER_DEFINED


def fortran_e(v, d=15, w=23):
    """One value in the Fortran edit descriptor Ew.d: 0.ddd...E+xx, right-justified in w."""
    v = float(v)
    if v == 0.0:
        s = "0." + "0" * d + "E+00"
    elif not np.isfinite(v):
        s = "NaN" if v != v else ("Infinity" if v > 0 else "-Infinity")
        return s.rjust(w)
    else:
        m, e = ("%.*e" % (d - 1, abs(v))).split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        exp = ("E%+03d" % ex) if abs(ex) <= 99 else ("%+04d" % ex)
        s = ("-" if v < 0 else "") + "0." + digits + exp
    return s.rjust(w)


def value_row(time, values):
    """WriteOnFile_Value: '(n(E23.15,1x))' over the time and the values (the blank after the
    last field is a position, not a character: the record ends with the last digit)."""
    return " ".join(fortran_e(x) for x in [time] + list(values)) + "\n"


class Probe:
    """One PROBE block.  components is None for 'All'."""

    def __init__(self, location, file, variable="", node=0, element=0, gauss_point=0, components=None,
                 nodes=None):
        self.location = location
        self.file = file
        self.variable = variable
        self.var_id = variable_id(variable)
        self.node = node
        self.element = element
        self.gauss_point = gauss_point
        self.components = components
        self.nodes = nodes
        self.active = True
        self.rows = []  # (time, values) or message, as written

    @property
    def needs_mesh(self):
        """Gauss-point and nodal-force probes read the replayed state of the mesh."""
        return self.location in ("gausspoint", "nodalforce")

    # ---- file
    def initialize(self, out_dir):
        with open(os.path.join(out_dir, self.file), "w") as f:
            f.write(HEADER)

    def _message(self, out_dir, text, deactivate=True):
        with open(os.path.join(out_dir, self.file), "a") as f:
            f.write(text + "\n")
        self.rows.append(text)
        if deactivate:
            self.active = False

    def _values(self, out_dir, time, values):
        values = np.asarray(values, dtype=np.float64).reshape(-1)
        with open(os.path.join(out_dir, self.file), "a") as f:
            f.write(value_row(time, values))
        self.rows.append((time, values.copy()))

    def _select(self, v):
        return v if self.components is None else np.asarray(v)[np.asarray(self.components) - 1]

    def _too_large(self, n):
        return self.components is not None and any(c > n for c in self.components)

    # ---- writers
    def write(self, out_dir, rec):
        if not self.active:
            return
        getattr(self, "_write_" + self.location)(out_dir, rec)

    def _write_node(self, out_dir, rec):
        if self.node < 1:
            raise ProbeError("probe %s: no Node" % self.file)
        if self.node > rec.nn:
            return self._message(out_dir, "ERROR - Node Number is greater than the Max Number of Nodes in the Mesh")
        if self.var_id != DISPLACEMENTS:
            return self._message(out_dir, "ERROR: Variable not available")
        u = rec.U[3 * (self.node - 1):3 * self.node]
        if self.components is None:
            return self._values(out_dir, rec.time, u)
        if self._too_large(3):
            return self._message(out_dir, "ERROR: Component is greater than the available")
        self._values(out_dir, rec.time, self._select(u))

    def _write_gausspoint(self, out_dir, rec):
        if self.element < 1 or self.gauss_point < 1:
            raise ProbeError("probe %s: no Element or Gauss Point" % self.file)
        if self.element > rec.ne:
            return self._message(out_dir, "ERROR - Number of Element is greater than Max Number of Elements in Mesh")
        if self.gauss_point > 8:
            return self._message(out_dir, "ERROR - Number of Gauss Point is greater than Max Number of Gauss Points "
                                          "in the Element")
        e, g = self.element - 1, self.gauss_point - 1
        if self.var_id == CAUCHY_STRESS:
            if self._too_large(6):
                return self._message(out_dir, "ERROR - Stress component is greater than Max Stress Size")
            return self._values(out_dir, rec.time, self._select(rec.stress(e, g)))
        if self.var_id == LOGARITHMIC_STRAIN:
            if self._too_large(6):
                return self._message(out_dir, "ERROR - Strain component is greater than Max Strain Size")
            return self._values(out_dir, rec.time, self._select(sym_voigt(log_strain(rec.F(e, g)))))
        if self.var_id == DEFORMATION_GRADIENT:
            if self._too_large(9):
                return self._message(out_dir, "ERROR - Deformation Gradient component is greater than Max Size")
            return self._values(out_dir, rec.time, self._select(voigt9(rec.F(e, g))))
        if self.var_id == USER_DEFINED:
            model, props = rec.material(e)
            for name, fn in user_results(model):
                if _key(name) == _key(self.variable):
                    v = np.atleast_1d(fn(rec, e, g, props))
                    if self._too_large(len(v)):
                        return self._message(out_dir, "ERROR - User defined variable component is greater than "
                                                      "Maximum.")
                    return self._values(out_dir, rec.time, self._select(v))
            return self._message(out_dir, "ERROR - User defined variable not found.")
        # the reference stops the program here (ModProbe.f90:558-559)
        raise ProbeError("Error in ModProbe - WriteProbeResult_GaussPoint - VariableNameID - not identified")

    def _write_nodalforce(self, out_dir, rec):
        if any(n > rec.nn for n in self.nodes):
            return self._message(out_dir, "ERROR - Node Number is greater than the Max Number of Nodes in the Mesh")
        fint = rec.internal_force()
        idx = 3 * (np.asarray(self.nodes) - 1)
        self._values(out_dir, rec.time, [fint[idx + d].sum() for d in range(3)])

    def _write_microstructure(self, out_dir, rec):
        if rec.homogenized is None:
            return self._message(out_dir, "Not a multiscale analysis")
        P9, F = rec.homogenized()
        v, err = microstructure_values(self.var_id, P9, F)
        if self._too_large(len(v)):
            return self._message(out_dir, err)
        self._values(out_dir, rec.time, self._select(v))


# ---- continuum mechanics of the probes (Tools/ModContinuumMechanics.f90, ModMathRoutines.f90)
def voigt9(A):
    """Convert_to_Voigt_3D (ModMathRoutines.f90:766-785): column by column, A11 A21 A31 A12 ..."""
    return np.asarray(A, dtype=np.float64).T.reshape(9)


def sym_voigt(A):
    """Convert_to_Voigt_3D_Sym / Tensor2ToVoigtSym: A11 A22 A33 A12 A23 A13."""
    A = np.asarray(A, dtype=np.float64)
    return np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1], A[1, 2], A[0, 2]])


def sym_tensor(v):
    """Convert_to_Tensor_3D_Sym."""
    return np.array([[v[0], v[3], v[5]], [v[3], v[1], v[4]], [v[5], v[4], v[2]]], dtype=np.float64)


def log_strain(F):
    """StrainMeasure(F, Logarithmic), ModContinuumMechanics.f90:30-101: sum over the eigenpairs
    (l_k, n_k) of b = F F^T of ln(sqrt(l_k)) n_k (x) n_k."""
    F = np.asarray(F, dtype=np.float64)
    lam, V = np.linalg.eigh(F @ F.T)
    return (V * np.log(np.sqrt(lam))) @ V.T


def almansi_strain(F):
    """StrainMeasure(F, Almansi): (I - (F F^T)^-1) / 2."""
    F = np.asarray(F, dtype=np.float64)
    return 0.5 * (np.eye(3) - np.linalg.inv(F @ F.T))


def von_mises(T):
    """vonMisesMeasure, ModContinuumMechanics.f90:177-191: sqrt(3 J2), J2 = dev T : dev T / 2."""
    T = np.asarray(T, dtype=np.float64)
    d = T - np.trace(T) / 3.0 * np.eye(3)
    return np.sqrt(1.5 * np.sum(d * d))


def first_piola_from_cauchy(F, sigma):
    """StressTransformation(F, sigma, Cauchy, FirstPiola): J sigma F^-T."""
    F = np.asarray(F, dtype=np.float64)
    return np.linalg.det(F) * sigma @ np.linalg.inv(F.T)


def cauchy_from_first_piola(F, P):
    """StressTransformation(F, P, FirstPiola, Cauchy): P F^T / J."""
    F = np.asarray(F, dtype=np.float64)
    return P @ F.T / np.linalg.det(F)


def microstructure_values(var_id, P9, F):
    """Values of a Micro Structure probe from the homogenized first Piola stress P9 (9 values,
    column by column, as HomogenizeStress returns it) and deformation gradient F (3x3)
    (ModProbe.f90:824-898); returns (values, message when a component is out of range)."""
    if var_id == FIRST_PIOLA_STRESS:
        return np.asarray(P9, dtype=np.float64), \
            "ERROR - Homogenized First Piola component is greater than Max Stress Size"
    if var_id == DEFORMATION_GRADIENT:
        return voigt9(F), "ERROR - Homogenized Deformation Gradient component is greater than Max Size"
    if var_id == CAUCHY_STRESS:
        P = np.asarray(P9, dtype=np.float64).reshape(3, 3).T  # VoigtToTensor2
        return sym_voigt(cauchy_from_first_piola(F, P)), \
            "ERROR - Homogenized Cauchy Stress component is greater than Max Stress Size"
    raise ProbeError("Error in ModProbe - WriteProbeResult_MicroStructure - VariableNameID - not identified")


# ---- user-defined results: GetResult of each material (ConstitutiveModelsLibrary/Mod*.f90),
# (name, function of (record, element, Gauss point, props)) in ID order
def _stress(rec, e, g, props):
    return rec.stress(e, g)


def _j2_strain(rec, e, g, props):
    F = rec.F(e, g)
    return np.array([F[0, 0] - 1.0, F[1, 1] - 1.0, F[2, 2] - 1.0, F[0, 1] + F[1, 0], F[1, 2] + F[2, 1],
                     F[0, 2] + F[2, 0]])


def _von_mises(rec, e, g, props):
    return von_mises(sym_tensor(rec.stress(e, g)))


def _almansi(rec, e, g, props):
    return sym_voigt(almansi_strain(rec.F(e, g)))


def _cnh_first_piola(rec, e, g, props):
    F = rec.F(e, g)
    a = voigt9(first_piola_from_cauchy(F, sym_tensor(rec.stress(e, g))))
    return np.array([a[0], a[4], 0.0, a[3], a[1], 0.0])


def _cnh_deformation_grad(rec, e, g, props):
    a = voigt9(rec.F(e, g))
    return np.array([a[0], a[4], 0.0, a[3], a[1], 0.0])


def _matrix_dissipation(rec, e, g, props):
    dv = rec.state(e, g)[9:18]
    return props[5] * float(np.dot(dv, dv))


USER_RESULTS = {
    1: [("Stress", _stress)],                                                        # ModGeneralizedHookesLaw .f90:580-611
    2: [("Stress", _stress), ("Strain", _j2_strain), ("von Mises Stress", _von_mises)],  # ModJ2Plasticity.f90:638-690
    3: [("Stress", _stress)],                                                        # ModNeoHookean.f90:628-660
    4: [("Stress", _stress)],                                                        # ModNeoHookeanQ1P0.f90:662-694
    5: [("Stress", _stress), ("Almansi Strain", _almansi), ("von Mises Cauchy Stress", _von_mises)],  # ModStVenantKirchhoff.f90:787-830
    7: [("First Piola Stress", _cnh_first_piola), ("Deformation Grad", _cnh_deformation_grad)],  # ModCompressibleNeoHookean.f90:559-610
    12: [("Matrix_Dissipation", _matrix_dissipation)],                               # ModViscoelasticMatrix.f90:1034-1060
}


def user_results(model):
    return USER_RESULTS.get(model, [])


# ---- input
def _lines(text):
    out = []
    for ln in text.splitlines():
        ln = ln.split("!", 1)[0].strip()
        if ln:
            out.append(ln)
    return out


def read_load_collector_nodes(path, collector):
    """Nodes of a load collector of a HyperMesh .cdb file, as NodalForceProbeConstructor reads
    them (ModProbe.f90:640-700): after the '!!hmname loadcol' line naming it and two more lines,
    the second field of every line of four comma-separated fields, in file order (a node listed
    for several DOFs appears as often)."""
    if not os.path.exists(path):
        raise ProbeError("HyperMesh file %s not found" % path)
    lines = open(path, encoding="latin-1").read().splitlines()
    i = 0
    while i < len(lines):
        if _key(lines[i]) == "!!hmnameloadcol" and i + 1 < len(lines):
            parts = lines[i + 1].split('"')
            if len(parts) > 1 and _key(parts[1]) == _key(collector):
                nodes = []
                j = i + 4
                while j < len(lines):
                    f = lines[j].split(",")
                    if len(f) != 4:
                        break
                    nodes.append(int(f[1]))
                    j += 1
                if not nodes:
                    raise ProbeError("Vector NodalForceList is empty")
                return nodes
        i += 1
    raise ProbeError("Load Collector (Nodal Force) could not be found in HyperMesh .cdb File")


def _components(text):
    if _key(text) == "all":
        return None
    parts = [p for p in text.split(",") if p.strip()]
    if not parts:
        raise ProbeError("probe without Components")
    return [int(float(p)) for p in parts]


def read_probes(path, case_dir=None):
    """The probes of a Post_Processing.post (ModExportResultFile.f90:140-240), in file order.
    Returns [] when the file does not exist."""
    if not os.path.exists(path):
        return []
    case_dir = case_dir or os.path.dirname(path)
    lines = _lines(open(path, encoding="latin-1").read())
    count = None
    for ln in lines:
        if "=" in ln and _key(ln.split("=", 1)[0]) == "numberofprobes":
            count = int(float(ln.split("=", 1)[1]))
            break
    probes, i = [], 0
    while i < len(lines) and (count is None or len(probes) < count):
        if _key(lines[i]) != "probe":
            i += 1
            continue
        opts = {}
        i += 1
        while i < len(lines) and _key(lines[i]) != "endprobe":
            if "=" not in lines[i]:
                raise ProbeError("Expression not identified in %s" % path)
            name, value = lines[i].split("=", 1)
            k = _OPTIONS.get(_key(name))
            if k is None:
                raise ProbeError("Expression not identified in %s" % path)
            opts[k] = value.strip()
            i += 1
        if i >= len(lines):
            raise ProbeError("Expecting Word END PROBE in %s" % path)
        i += 1
        loc = _key(opts.get("location", ""))
        if loc not in LOCATIONS:
            raise ProbeError("probe location %r not identified" % opts.get("location", ""))
        file = opts.get("file", "")
        if loc == "nodalforce":
            hm = os.path.join(case_dir, opts.get("hypermesh_file", ""))
            probes.append(Probe(loc, file, nodes=read_load_collector_nodes(hm, opts.get("load_collector", ""))))
            continue
        comps = _components(opts.get("components", ""))
        probes.append(Probe(loc, file, variable=opts.get("variable", ""), node=int(float(opts.get("node", 0))),
                            element=int(float(opts.get("element", 0))),
                            gauss_point=int(float(opts.get("gauss_point", 0))), components=comps))
    if count is not None and len(probes) < count:
        raise ProbeError("Expecting Word PROBE in %s" % path)
    return probes


class Record:
    """What the probes read at one finished step of the replay: time, U, and lazily the
    Gauss-point fields of the replayed mesh (stress (ne, 8, 6), F (ne, 8, 3, 3), state rows),
    its internal force and the element materials."""

    def __init__(self, time, U, nn, ne, mesh=None, materials=None, mat_of_element=None, homogenized=None):
        self.time, self.U, self.nn, self.ne = float(time), U, nn, ne
        self._mesh = mesh
        self._materials = materials or {}
        self._mat = mat_of_element
        self.homogenized = homogenized
        self._sig = self._F = self._state = self._fint = None

    def stress(self, e, g):
        if self._sig is None:
            self._sig = self._mesh.gp_stress()
        return self._sig[e, g]

    def F(self, e, g):
        if self._F is None:
            self._F = self._mesh.gp_deformation_gradient(self.U)
        return self._F[e, g]

    def state(self, e, g):
        if self._state is None:
            self._state = self._mesh.state(1)
        return self._state[e, g]

    def internal_force(self):
        if self._fint is None:
            self._fint, _ = self._mesh.internal_force()
        return self._fint

    def material(self, e):
        mid = int(self._mat[e]) if self._mat is not None else 1
        return self._materials[mid]


def probe_file_rows(path):
    """(header, rows) of a probe file: rows are (time, values) or message strings."""
    with open(path) as f:
        lines = f.read().splitlines()
    rows = []
    num = re.compile(r"^\s*[-+0-9.]")
    for ln in lines[1:]:
        if num.match(ln):
            v = [float(x) for x in ln.split()]
            rows.append((v[0], np.array(v[1:])))
        elif ln.strip():
            rows.append(ln)
    return lines[0], rows
