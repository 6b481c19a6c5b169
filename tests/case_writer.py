"""Writers of the reference's input files, for tests that build a case in a temporary directory:
a HyperMesh ANSYS deck in the layout ReadMeshHyperMesh / ReadBoundaryConditionsHyperMesh read
(FEM/Input/ModReadInputFile.f90:909-1395: NUMOFF, NBLOCK with its format line, ET 185, EBLOCK
with its format line and 19 fields per hex8, '-1', the load-step data with one '!!HMNAME LOADCOL'
collector per node set), and the Settings / time / table / probe files.  Test infrastructure."""
import os


def write_cdb(path, X, conn, mat, collectors):
    """collectors: [(name, [(node, dof letter pair e.g. 'UX')...])], nodes 1-based."""
    with open(path, "w") as f:
        f.write("!! ANSYS Input Deck written for a test\n/PREP7\n")
        f.write("NUMOFF,NODE,%8d\nNUMOFF,ELEM,%8d\n" % (len(X), len(conn)))
        f.write("NBLOCK,6,SOLID\n(3i8,6e16.9)\n")
        for i, x in enumerate(X, start=1):
            f.write("%8d%8d%8d%16.9E%16.9E%16.9E\n" % (i, 0, 0, x[0], x[1], x[2]))
        f.write("N,R5.3,LOC,     -1,\nET,1,185\nEBLOCK,19,SOLID,\n(19i8)\n")
        for e, (c, m) in enumerate(zip(conn, mat), start=1):
            f.write("".join("%8d" % v for v in [m, 1, 1, 0, 0, 0, 0, 0, 8, 0, e] + list(c)) + "\n")
        f.write("      -1\n!! Loadstep Data\n")
        for k, (name, entries) in enumerate(collectors, start=1):
            f.write('!!HMNAME LOADCOL\n!!%10d "%s"\n!!HWCOLOR LOADCOL\n!!%10d        5\n' % (k, name, k))
            for node, dof in entries:
                f.write("D,%d,%s,             0.0\n" % (node, dof))
        f.write("LSWRITE,1\nFINISH\n")


def write_text(directory, name, text):
    with open(os.path.join(directory, name), "w") as f:
        f.write(text)


def write_table(directory, name, rows):
    write_text(directory, name, "Time\tValue\n" + "".join("%.17g\t%.17g\n" % r for r in rows))


def write_time(directory, cases, name="Time_Discretization.dat"):
    """cases: [(t0, t1, steps)]."""
    txt = "Number of Load Cases\n%d\nLoad Case\tInitial Time\tFinal Time\tSteps\n" % len(cases)
    txt += "".join("%d\t%.17g\t%.17g\t%d\n" % (k, a, b, n) for k, (a, b, n) in enumerate(cases, start=1))
    write_text(directory, name, txt)


SETTINGS_MULTISCALE = """ANALYSIS SETTINGS
    Problem Type = Mechanical
    Analysis Type = Quasi Static
    Nonlinear Analysis = True
    Hypothesis of Analysis = 3D
    Element Technology = Full Integration
    Maximum Cut Backs = 10
    Multiscale Analysis = True
    Multiscale Model = Periodic
    Restart = False
END ANALYSIS SETTINGS

LINEAR SOLVER
    PARDISO
END LINEAR SOLVER

NONLINEAR SOLVER
    Newton_Raphson_Full
    Tol = 1e-10
    MaxIter = 12
END NONLINEAR SOLVER

MATERIAL
    Number of Materials = 2
    Material ID = 1
    Neo_Hookean
    C10 = 0.5
    Bulk Modulus = 50.0
    Material ID = 2
    Neo_Hookean
    C10 = 5.0
    Bulk Modulus = 500.0
END MATERIAL

MESH AND BOUNDARY CONDITIONS
    Mesh File = Mesh.cdb
    Preprocessor = HyperMesh
    Time Discretization File = Time_Discretization.dat
END MESH AND BOUNDARY CONDITIONS

MACROSCOPIC DEFORMATION GRADIENT
    F11 = F11
    F12 = F12.tab
    F13 = Zero
    F21 = Zero
    F22 = One
    F23 = Zero
    F31 = Zero
    F32 = Zero
    F33 = F33
END MACROSCOPIC DEFORMATION GRADIENT
"""

POST_MULTISCALE = """POST PROCESSOR
    Post Processor = None
END POST PROCESSOR

NUMBER OF PROBES = 4

PROBE
    Location = Micro Structure
    File Name = P.probe
    Variable Name = First Piola Stress
    Components = All
END PROBE

PROBE
    Location = Micro Structure
    File Name = F.probe
    Variable Name = Deformation Gradient
    Components = 1,4,9
END PROBE

PROBE
    Location = Micro Structure
    File Name = S.probe
    Variable Name = Cauchy Stress
    Components = All
END PROBE

PROBE
    Location = Gauss Point
    File Name = gp.probe
    Variable Name = Cauchy Stress
    Element = 5
    Gauss Point = 3
    Components = All
END PROBE
"""

MACRO_TABLES = {"F11": [(0.0, 1.0), (1.0, 1.10)], "F12.tab": [(0.0, 0.0), (1.0, 0.05)],
                "F33": [(0.0, 1.0), (1.0, 0.95)]}


def write_multiscale_case(directory, X, conn, mat, steps=3):
    """A periodic multiscale case of the RVE mesh (X, conn, mat ids 1/2): its boundary node set as
    the 'boundary' collector, the macro F of MACRO_TABLES over one load case of `steps` steps."""
    import numpy as np
    lo, hi = X.min(axis=0), X.max(axis=0)
    on = np.any((np.abs(X - lo) < 1e-9) | (np.abs(X - hi) < 1e-9), axis=1)
    bnd = [(int(a) + 1, "UX") for a in np.nonzero(on)[0]]
    write_cdb(os.path.join(directory, "Mesh.cdb"), X, conn, mat, [("boundary", bnd)])
    write_text(directory, "Settings.dat", SETTINGS_MULTISCALE)
    write_text(directory, "Post_Processing.post", POST_MULTISCALE)
    write_time(directory, [(0.0, 1.0, steps)])
    write_table(directory, "F11.tab", MACRO_TABLES["F11"])
    write_table(directory, "F12.tab", MACRO_TABLES["F12.tab"])
    write_table(directory, "F33.tab", MACRO_TABLES["F33"])
