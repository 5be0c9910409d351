"""ompVC parameter decks: 97 positional ``NAME = VALUE`` entries.

Entry order follows ompVC/parameters.c:30-138; the reader there ignores names,
so order is the format.  ``write_deck`` emits every entry; ``read_deck`` parses
one back (names are checked only to detect 96-entry decks without Num_threads).
"""

KEYS = [
    ("Path_Data", "s", "pwd"), ("Path_Output", "s", "pwd"), ("Dimensions", "d", 3),
    ("Data_MeshType", "d", 1), ("Data_InFilePrefix", "s", "data"), ("Data_SuffixTMin", "d", 0),
    ("Data_SuffixTDelta", "d", 1), ("Data_TRes", "d", 2), ("Data_TDelta", "f", 1.0),
    ("Data_TMin", "f", 0.0), ("Data_TPeriodic", "d", 0),
    ("Data_MeshBounds.XMin", "f", 0.0), ("Data_MeshBounds.XMax", "f", 1.0),
    ("Data_MeshBounds.YMin", "f", 0.0), ("Data_MeshBounds.YMax", "f", 1.0),
    ("Data_MeshBounds.ZMin", "f", 0.0), ("Data_MeshBounds.ZMax", "f", 1.0),
    ("Fluid_Density", "f", 1.0), ("Fluid_Viscosity", "f", 1.0),
    ("Output_TStart", "f", 0.0), ("Output_TRes", "d", 1), ("Output_TDelta", "f", 1.0),
    ("Int_Type", "d", 1), ("Int_TimeStep", "f", 0.001), ("Int_Accuracy", "f", 0.0001),
    ("Int_MinTimeStep", "f", 0.0001), ("Int_MaxTimeStep", "f", 0.001), ("Int_TimeDirection", "d", 1),
    ("Int_NormalFlow", "d", 0), ("Int_NormalFlowScaling", "f", 1.0), ("Int_Extrapolate", "d", 0),
    ("Particle_Radius", "f", 0.0), ("Particle_Density", "f", 1.0), ("Particle_ICType", "d", 0),
    ("Gravity_Vector[0]", "f", 0.0), ("Gravity_Vector[1]", "f", 0.0), ("Gravity_Vector[2]", "f", 0.0),
    ("LocalSearchChecking", "d", 0), ("Num_threads", "d", 0),
    ("FTLE_Compute", "d", 0), ("FTLE_GenerateMesh", "d", 1), ("FTLE_ICFile", "s", "ftle_ic.bin"),
    ("FTLE_MeshBounds.XMin", "f", 0.0), ("FTLE_MeshBounds.XMax", "f", 1.0),
    ("FTLE_MeshBounds.YMin", "f", 0.0), ("FTLE_MeshBounds.YMax", "f", 1.0),
    ("FTLE_MeshBounds.ZMin", "f", 0.0), ("FTLE_MeshBounds.ZMax", "f", 0.0),
    ("FTLE_MeshBounds.XRes", "d", 2), ("FTLE_MeshBounds.YRes", "d", 2), ("FTLE_MeshBounds.ZRes", "d", 1),
    ("FTLE_IntTLength", "f", 1.0), ("FTLE_ComputeVariation", "d", 0), ("FTLE_VariationOutFreq", "d", 1),
    ("FTLE_OutFilePrefix", "s", "ftle"),
    ("Trace_Compute", "d", 0), ("Trace_ReleaseStrategy", "d", 0), ("Trace_ReleaseTMax", "f", 1.0),
    ("Trace_GenerateMesh", "d", 0), ("Trace_InFile", "s", "seeds.bin"), ("Trace_MultipleInFiles", "d", 0),
    ("Trace_InFileFormat", "d", 4), ("Trace_OutFilePrefix", "s", "tracer"), ("Trace_NumLaunchTimes", "d", 1),
    ("Trace_LaunchTimeSpacing", "f", 1.0), ("Trace_IntTLength", "f", 1.0), ("Trace_AlwaysOutput", "d", 0),
    ("Trace_MeshBounds.XMin", "f", 0.0), ("Trace_MeshBounds.XMax", "f", 1.0),
    ("Trace_MeshBounds.YMin", "f", 0.0), ("Trace_MeshBounds.YMax", "f", 1.0),
    ("Trace_MeshBounds.ZMin", "f", 0.0), ("Trace_MeshBounds.ZMax", "f", 0.0),
    ("Trace_MeshBounds.XRes", "d", 1), ("Trace_MeshBounds.YRes", "d", 1), ("Trace_MeshBounds.ZRes", "d", 1),
    ("Trace_APCompute", "d", 0), ("Trace_CETCompute", "d", 0), ("Trace_CETAuxillaryMesh", "d", 0),
    ("Trace_CETMeshPrefix", "s", "aux"), ("Trace_CETSubsteps", "d", 2), ("Trace_RTCompute", "d", 0),
    ("Trace_RTOutFilePrefix", "s", "rt"),
    ("VelOut_Compute", "d", 0), ("VelOut_GenerateMesh", "d", 1), ("VelOut_InFile", "s", "velin.txt"),
    ("VelOut_InFileFormat", "d", 1), ("VelOut_FilePrefix", "s", "velout"),
    ("VelOut_MeshBounds.XMin", "f", 0.0), ("VelOut_MeshBounds.XMax", "f", 1.0),
    ("VelOut_MeshBounds.YMin", "f", 0.0), ("VelOut_MeshBounds.YMax", "f", 1.0),
    ("VelOut_MeshBounds.ZMin", "f", 0.0), ("VelOut_MeshBounds.ZMax", "f", 0.0),
    ("VelOut_MeshBounds.XRes", "d", 1), ("VelOut_MeshBounds.YRes", "d", 1), ("VelOut_MeshBounds.ZRes", "d", 1),
]
assert len(KEYS) == 97
_NAMES = [k for k, _, _ in KEYS]


def _fmt(t, v):
    if t == "s":
        return str(v)
    if t == "d":
        return "%d" % int(v)
    return repr(float(v))


def write_deck(path, **values):
    """Write a 97-entry deck; keyword names are the entry names with '.' and
    '[...]' replaced by '_' (e.g. FTLE_MeshBounds_XRes, Gravity_Vector_0)."""
    norm = {k.replace(".", "_").replace("[", "_").replace("]", ""): k for k in _NAMES}
    vals = {}
    for key, v in values.items():
        if key not in norm:
            raise KeyError(key)
        vals[norm[key]] = v
    with open(path, "w") as f:
        f.write("# flowVC deck: 97 positional entries (ompVC/parameters.c:30-138)\n")
        for name, t, dflt in KEYS:
            f.write("%s = %s\n" % (name, _fmt(t, vals.get(name, dflt))))


def read_deck(path):
    """Parse a deck into {name: value} by position.  As ompVC/parameters.c:258-292,
    a '#' where an entry name is expected starts a comment to end of line."""
    text = open(path).read()
    pos, n = 0, len(text)

    def token(allow_comment):
        nonlocal pos
        while True:
            while pos < n and text[pos].isspace():
                pos += 1
            if pos >= n:
                return None
            if allow_comment and text[pos] == "#":
                while pos < n and text[pos] != "\n":
                    pos += 1
                continue
            start = pos
            while pos < n and not text[pos].isspace():
                pos += 1
            return text[start:pos]

    out, k = {}, 0
    while k < 97:
        name = token(True)
        if name is None:
            break
        if k == 38 and name == "FTLE_Compute":
            out["Num_threads"] = 0
            k += 1
        if token(False) is None:
            break
        v = token(False)
        if v is None:
            break
        kname, typ, _ = KEYS[k]
        out[kname] = v if typ == "s" else (int(float(v)) if typ == "d" else float(v))
        k += 1
    if k < 97:
        raise ValueError("deck %s has %d entries" % (path, k))
    return out
