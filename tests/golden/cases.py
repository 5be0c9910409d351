"""Golden cases shared by make_goldens.py (writes fixtures by running the
reference ompVC built from /root/reference in the authoring container) and the
tests (regenerate the same inputs, run the oracle or the GPU driver, compare).

Inputs are regenerated deterministically from the product's generators; their
sha256 digests are recorded in each case's manifest so a generator change is
caught before any comparison.
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "flowvc-1_amd"))

from fvcpy import deck, host  # noqa: E402

DG = dict(Dimensions=2, Data_MeshType=0, Data_InFilePrefix="dg", Data_TRes=21, Data_TDelta=0.5, Data_TMin=0.0,
          Data_TPeriodic=0, Data_MeshBounds_XMin=0.0, Data_MeshBounds_XMax=2.0, Data_MeshBounds_YMin=0.0,
          Data_MeshBounds_YMax=1.0, Data_MeshBounds_ZMin=0.0, Data_MeshBounds_ZMax=0.0, Int_Type=1,
          Int_TimeStep=0.01)
PIPE = dict(Dimensions=3, Data_MeshType=1, Data_InFilePrefix="pipe", Data_TRes=51, Data_TDelta=0.02, Data_TMin=0.0,
            Data_TPeriodic=1, Data_MeshBounds_XMin=-1.0, Data_MeshBounds_XMax=1.0, Data_MeshBounds_YMin=-1.0,
            Data_MeshBounds_YMax=1.0, Data_MeshBounds_ZMin=0.0, Data_MeshBounds_ZMax=10.0, Int_Type=1,
            Int_TimeStep=0.001, LocalSearchChecking=1)


def gen_dg(d):
    """Double gyre A=0.1, eps=0.25, omega=2 pi/10 on 64x32 nodes, 21 slices t=0..10."""
    host.write_double_gyre(d, "dg", 64, 32, 21, 0.5)


PIPE_DIMS = (12, 12, 40)


def gen_pipe(d):
    """Kuhn-split pipe R=1 L=10 on 12x12x40 voxels; 50 noisy slices (sigma 0.01, seed 1234)."""
    t = host.pipe(*PIPE_DIMS, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    host.write_unstruct_mesh(d, "pipe", xyz, conn, adj)
    for s in range(50):
        host.write_vel_slice(os.path.join(d, "pipe_vel.%d.bin" % s), s * 0.02,
                             host.tube_velocity(t, xyz, s * 0.02, 0.01, 1234, s))
    host.write_binary_list(os.path.join(d, "seeds.bin"), host.tube_seeds(t, 1000, 0.0, 10.0 / 3.0, 42))
    # location probes: interior, face, edge, vertex, outside and far-apart points
    pts = list(host.tube_seeds(t, 40, 0.0, 10.0, 7))
    dx, dz = 2.0 / PIPE_DIMS[0], 10.0 / PIPE_DIMS[2]
    for i in range(2, 10):
        x = -1.0 + i * dx
        pts.append([x, x, 3 * dz])                      # on Kuhn diagonal faces (x=y)
        pts.append([x, -1.0 + 5 * dx, 7 * dz + 0.3 * dz])  # on voxel face planes
        pts.append([x, -1.0 + 6 * dx, 11 * dz])          # on voxel edges
        pts.append([-1.0 + 5 * dx, -1.0 + 6 * dx, i * dz])  # on mesh vertices
    pts += [[0.0, 0.0, -0.5], [1.5, 0.0, 5.0], [0.0, 0.0, 10.5], [0.99, 0.99, 2.0], [0.0, 0.0, 0.0],
            [0.0, 0.0, 10.0], [-1.0 + 6 * dx, 0.0, 0.0]]
    with open(os.path.join(d, "probes.txt"), "w") as f:
        f.write("%d\n" % len(pts))
        for p in pts:
            f.write("%.17g %.17g %.17g\n" % tuple(p))


def release_line():
    """21 release points across the pipe at z = 0.5, 0.09 apart (the spacing
    that sets the release interval)."""
    return np.array([[-0.9 + 0.09 * i, 0.013, 0.5] for i in range(21)])


def gen_pipe_stagger(d):
    gen_pipe(d)
    host.write_binary_list(os.path.join(d, "release.bin"), release_line())


def gen_pipe_steady(d):
    """The pipe with every slice equal to the t=0 field (steady flow: linear
    release-time solves)."""
    t = host.pipe(*PIPE_DIMS, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    host.write_unstruct_mesh(d, "pipe", xyz, conn, adj)
    v0 = host.tube_velocity(t, xyz, 0.0, 0.01, 1234, 0)
    for s in range(50):
        host.write_vel_slice(os.path.join(d, "pipe_vel.%d.bin" % s), s * 0.02, v0)
    host.write_binary_list(os.path.join(d, "release.bin"), release_line())


def strain_pipe(ne, s):
    """Per-element strain rate of frame s (any field serves the integration's parity)."""
    e = np.arange(ne, dtype=np.float64)
    return 0.5 + 0.4 * np.sin(0.001 * e + 0.3 * s)


def gen_pipe_ap(d):
    gen_pipe(d)
    ne = int(np.fromfile(os.path.join(d, "pipe_connectivity.bin"), np.int32, count=1)[0])
    for s in range(50):
        with open(os.path.join(d, "pipe_strain-rate.%d.bin" % s), "wb") as f:
            np.array([s * 0.02]).tofile(f)
            strain_pipe(ne, s).tofile(f)


def gen_dg_ap(d):
    gen_dg(d)
    i = np.arange(63 * 31, dtype=np.float64)
    for s in range(21):
        with open(os.path.join(d, "dg_strain-rate.%d.bin" % s), "wb") as f:
            np.array([s * 0.5]).tofile(f)
            (1.0 + 0.5 * np.cos(0.05 * i + 0.3 * s)).tofile(f)


def velout_points():
    """Listed velocity output points: interior, on voxel faces and vertices,
    and outside the mesh."""
    t = host.pipe(*PIPE_DIMS, 1.0, 10.0)
    pts = list(host.tube_seeds(t, 60, 0.0, 10.0, 11))
    dx, dz = 2.0 / PIPE_DIMS[0], 10.0 / PIPE_DIMS[2]
    for i in range(2, 10):
        pts.append([-1.0 + i * dx, -1.0 + 5 * dx, 7 * dz])
        pts.append([-1.0 + 5 * dx, -1.0 + i * dx, 13 * dz + 0.37 * dz])
    pts += [[0.99, 0.99, 2.0], [0.0, 0.0, 10.5], [0.0, 0.0, 0.0]]
    return np.array(pts)


def gen_pipe_velout(d):
    gen_pipe(d)
    with open(os.path.join(d, "velpts.txt"), "w") as f:
        p = velout_points()
        f.write("%d\n" % p.shape[0])
        for q in p:
            f.write("%.17g %.17g %.17g\n" % tuple(q))


def gen_pipe_multi(d):
    """The pipe with one binary seed file per release (Trace_MultipleInFiles,
    tracers.c:1007-1016): 300, 500 and 200 seeds."""
    gen_pipe(d)
    t = host.pipe(*PIPE_DIMS, 1.0, 10.0)
    for ff, (cnt, seed) in enumerate(((300, 5), (500, 6), (200, 7))):
        host.write_binary_list(os.path.join(d, "mseeds.%d.bin" % ff), host.tube_seeds(t, cnt, 0.0, 10.0 / 3.0, seed))


def pipe_walls():
    """Wall nodes of the pipe (on a boundary face, strictly between inlet and
    outlet) and their inward normals, radial towards the axis."""
    t = host.pipe(*PIPE_DIMS, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    opposite = [1, 2, 3, 0]  # adjacency slot -> node opposite the face it crosses
    nodes = set()
    for s in range(4):
        e = np.nonzero(adj[:, s] < 0)[0]
        keep = [k for k in range(4) if k != opposite[s]]
        nodes.update(conn[e][:, keep].ravel().tolist())
    ids = np.array(sorted(nodes), np.int32)
    z = xyz[ids, 2]
    ids = ids[(z > 1e-9) & (z < 10.0 - 1e-9)]
    r = np.hypot(xyz[ids, 0], xyz[ids, 1])
    nrm = np.stack([-xyz[ids, 0] / r, -xyz[ids, 1] / r, np.zeros(ids.size)], 1)
    return ids, nrm


def gen_pipe_normal(d):
    """The pipe with <prefix>_normals.bin (mesh.c:255-300): count, node ids,
    inward normals."""
    gen_pipe(d)
    write_normals(d)


def gen_pipe_stagger_normal(d):
    gen_pipe_stagger(d)
    write_normals(d)


def write_normals(d):
    ids, nrm = pipe_walls()
    with open(os.path.join(d, "pipe_normals.bin"), "wb") as f:
        np.array([ids.size], "<i4").tofile(f)
        ids.astype("<i4").tofile(f)
        np.ascontiguousarray(nrm, "<f8").tofile(f)


def gen_dg_velout(d):
    gen_dg(d)
    rng = np.random.default_rng(9)
    p = np.concatenate([rng.uniform([-0.1, -0.1], [2.1, 1.1], size=(200, 2)), [[0.0, 0.0], [2.0, 1.0], [1.0, 0.5]]])
    with open(os.path.join(d, "velpts.txt"), "w") as f:
        f.write("%d\n" % p.shape[0])
        for q in p:
            f.write("%.17g %.17g\n" % tuple(q))


def write_cart_header(path, mn, mx, res):
    """Cartesian mesh header: XMin, XMax (double), XRes (int), then y, then z
    (mesh.c:67-122)."""
    with open(path, "wb") as f:
        for a in range(3):
            np.array([mn[a], mx[a]], "<f8").tofile(f)
            np.array([res[a]], "<i4").tofile(f)


BOX_RES = (25, 13, 13)


def box_velocity(t):
    """3-D unsteady field on [0,2]x[0,1]^2 (a double gyre in x-y, stretched and
    sheared in z), rows in k, j, i order with i fastest (velocity.c:129-140).
    It leaves through z=1 where cos(pi x) > 0 and through x at the z-shear."""
    nx, ny, nz = BOX_RES
    z, y, x = np.meshgrid(np.linspace(0.0, 1.0, nz), np.linspace(0.0, 1.0, ny), np.linspace(0.0, 2.0, nx),
                          indexing="ij")
    A, eps, om = 0.1, 0.25, 2 * np.pi / 10
    a, b = eps * np.sin(om * t), 1 - 2 * eps * np.sin(om * t)
    f, df = a * x * x + b * x, 2 * a * x + b
    u = -np.pi * A * np.sin(np.pi * f) * np.cos(np.pi * y) + 0.05 * np.cos(np.pi * z)
    v = np.pi * A * np.cos(np.pi * f) * np.sin(np.pi * y) * df
    w = 0.1 * np.cos(np.pi * x) + 0.05 * np.sin(2 * np.pi * z) * (1 + 0.5 * np.sin(om * t))
    return np.stack([u.ravel(), v.ravel(), w.ravel()], 1)


def gen_box(d):
    write_cart_header(os.path.join(d, "box_Cartesian.bin"), [0.0, 0.0, 0.0], [2.0, 1.0, 1.0], BOX_RES)
    for s in range(11):
        host.write_vel_slice(os.path.join(d, "box_vel.%d.bin" % s), 0.5 * s, box_velocity(0.5 * s))


CHAN_CELLS = (32, 8)


def chan_mesh():
    """2-D channel [0,4]x[0,1] of jittered triangles (two per cell, diagonals
    alternating).  Slot order of the adjacency (mesh.c:593-600): [0] across the
    edge opposite node 1, [1] opposite node 2, [2] opposite node 0."""
    nx, ny = CHAN_CELLS
    X, Y = np.meshgrid(np.linspace(0.0, 4.0, nx + 1), np.linspace(0.0, 1.0, ny + 1), indexing="ij")
    rng = np.random.default_rng(21)
    inner = (np.arange(nx + 1)[:, None] % nx != 0) & (np.arange(ny + 1)[None, :] % ny != 0)
    X = X + inner * rng.uniform(-0.2, 0.2, X.shape) * (4.0 / nx)
    Y = Y + inner * rng.uniform(-0.2, 0.2, Y.shape) * (1.0 / ny)
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    tris = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, e = i * (ny + 1) + j, (i + 1) * (ny + 1) + j, (i + 1) * (ny + 1) + j + 1, i * (ny + 1) + j + 1
            tris += [[a, b, c], [a, c, e]] if (i + j) % 2 == 0 else [[a, b, e], [b, c, e]]
    edges = {}
    for k, (n0, n1, n2) in enumerate(tris):
        for slot, (p, q) in enumerate(((n0, n2), (n0, n1), (n1, n2))):
            edges.setdefault((min(p, q), max(p, q)), []).append((k, slot))
    adj = -np.ones((len(tris), 4), np.int32)
    for lst in edges.values():
        if len(lst) == 2:
            (k1, s1), (k2, s2) = lst
            adj[k1, s1], adj[k2, s2] = k2, k1
    conn = np.concatenate([np.array(tris), np.zeros((len(tris), 1), np.int64)], 1).astype(np.int32)
    return xyz, conn, adj


def chan_velocity(xyz, t):
    """Pulsatile channel flow with a cross-stream wave (zero on the walls)."""
    x, y = xyz[:, 0], xyz[:, 1]
    u = 6.0 * y * (1.0 - y) * (1.0 + 0.5 * np.sin(2 * np.pi * t))
    v = 0.1 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(2 * np.pi * t)
    return np.stack([u, v, np.zeros_like(u)], 1)


def gen_chan(d):
    xyz, conn, adj = chan_mesh()
    host.write_unstruct_mesh(d, "chan", xyz, conn, adj)
    for s in range(21):
        host.write_vel_slice(os.path.join(d, "chan_vel.%d.bin" % s), 0.05 * s, chan_velocity(xyz, 0.05 * s))
    rng = np.random.default_rng(17)
    pts = rng.uniform([0.1, 0.01], [3.9, 0.99], size=(600, 2))
    # nodes and edge midpoints: points held by several triangles, where the chain's guesses decide
    pts = np.concatenate([pts, xyz[::7, :2], 0.5 * (xyz[conn[::5, 0], :2] + xyz[conn[::5, 1], :2]),
                          [[4.5, 0.5], [-0.1, 0.5], [2.0, 1.2]]])
    host.write_binary_list(os.path.join(d, "seeds.bin"), pts, dims=2)


STEN_DATA = os.path.join(HERE, "stenosis2d_data")
STEN_SLICES = (7070, 7080, 7090, 7100)


def sten_mesh():
    """The reference's 2-D stenosis mesh (cudaVC/data/Stenosis2D/bin: 40,279
    nodes, 79,066 triangles, adjacency with a 0 in the unused fourth slot)."""
    xyz = np.fromfile(os.path.join(STEN_DATA, "sineflow_coordinates.bin"), "<f8", offset=4).reshape(-1, 3)
    conn = np.fromfile(os.path.join(STEN_DATA, "sineflow_connectivity.bin"), "<i4", offset=4).reshape(-1, 4)
    return xyz, conn


def gen_sten(d):
    """The reference's own Stenosis2D data, copied unchanged: mesh files and
    the first four velocity slices (sineflow_vel.7070 ... 7100, t = 1.414 ...
    1.420 s); the runs below read no others."""
    for f in ["sineflow_coordinates.bin", "sineflow_connectivity.bin", "sineflow_adjacency.bin"] + \
            ["sineflow_vel.%d.bin" % k for k in STEN_SLICES]:
        shutil.copy(os.path.join(STEN_DATA, f), os.path.join(d, f))


def sten_probes():
    """Location probes on the stenosis: random points over the bounding box
    (the walls curve through it, so some fall outside), mesh vertices, edge
    midpoints (points held by two or more triangles, where the chain's guesses
    decide) and far-apart points."""
    xyz, conn = sten_mesh()
    rng = np.random.default_rng(31)
    pts = [rng.uniform([-0.0171, 0.0], [0.0171, 0.00275], size=(300, 2)), xyz[::211, :2],
           0.5 * (xyz[conn[::157, 0], :2] + xyz[conn[::157, 1], :2]),
           [[-0.0171, 0.0001], [0.0171, 0.0027], [0.0, 0.0014], [0.0, -0.001], [0.02, 0.001]]]
    return np.concatenate(pts)


def gen_sten_probes(d):
    gen_sten(d)
    p = sten_probes()
    with open(os.path.join(d, "probes.txt"), "w") as f:
        f.write("%d\n" % p.shape[0])
        for q in p:
            f.write("%.17g %.17g\n" % tuple(q))


def gen_sten_velout(d):
    gen_sten(d)
    p = sten_probes()[::2]
    with open(os.path.join(d, "velpts.txt"), "w") as f:
        f.write("%d\n" % p.shape[0])
        for q in p:
            f.write("%.17g %.17g\n" % tuple(q))


def gen_sten_seeds(d):
    """1,200 seeds over the upstream half of the channel and the throat."""
    gen_sten(d)
    rng = np.random.default_rng(37)
    host.write_binary_list(os.path.join(d, "seeds.bin"), rng.uniform([-0.012, 0.0002], [0.002, 0.0026], size=(1200, 2)),
                           dims=2)


BOX = dict(Dimensions=3, Data_MeshType=0, Data_InFilePrefix="box", Data_TRes=11, Data_TDelta=0.5, Data_TMin=0.0,
           Data_TPeriodic=0, Data_MeshBounds_XMin=0.0, Data_MeshBounds_XMax=2.0, Data_MeshBounds_YMin=0.0,
           Data_MeshBounds_YMax=1.0, Data_MeshBounds_ZMin=0.0, Data_MeshBounds_ZMax=1.0, Int_Type=1,
           Int_TimeStep=0.01)
STEN = dict(Dimensions=2, Data_MeshType=1, Data_InFilePrefix="sineflow", Data_SuffixTMin=7070, Data_SuffixTDelta=10,
            Data_TRes=101, Data_TDelta=0.002, Data_TMin=1.414, Data_TPeriodic=1, Data_MeshBounds_XMin=-0.01715,
            Data_MeshBounds_XMax=0.01715, Data_MeshBounds_YMin=0.0, Data_MeshBounds_YMax=0.00275,
            Data_MeshBounds_ZMin=0.0, Data_MeshBounds_ZMax=0.0, Int_Type=1, Int_TimeStep=0.00007,
            LocalSearchChecking=1)
CHAN = dict(Dimensions=2, Data_MeshType=1, Data_InFilePrefix="chan", Data_TRes=21, Data_TDelta=0.05, Data_TMin=0.0,
            Data_TPeriodic=0, Data_MeshBounds_XMin=0.0, Data_MeshBounds_XMax=4.0, Data_MeshBounds_YMin=0.0,
            Data_MeshBounds_YMax=1.0, Data_MeshBounds_ZMin=0.0, Data_MeshBounds_ZMax=0.0, Int_Type=1,
            Int_TimeStep=0.005, LocalSearchChecking=1)

VELOUT = {"VelOut_Compute": 1, "VelOut_MeshBounds_XMin": -1.0, "VelOut_MeshBounds_XMax": 1.0,
          "VelOut_MeshBounds_YMin": -1.0, "VelOut_MeshBounds_YMax": 1.0, "VelOut_MeshBounds_ZMin": 0.0,
          "VelOut_MeshBounds_ZMax": 10.0}

STAGGER = dict(PIPE, Num_threads=1, Output_TStart=0.0, Output_TRes=9, Output_TDelta=0.05, Int_TimeDirection=1,
               Trace_Compute=1, Trace_ReleaseStrategy=1, Trace_ReleaseTMax=0.3, Trace_GenerateMesh=0,
               Trace_InFile="release.bin", Trace_InFileFormat=4, Trace_IntTLength=0.25)

CASES = {
    # 2-D Cartesian FTLE over T=10 (the CPU-runnable workload's physics, 128x64 grid)
    "dg_ftle": dict(gen=gen_dg, outputs=["dgftle.0.bin", "dgftle_noT.0.bin", "dgftle_Cartesian.bin"],
                    deck=dict(DG, Output_TStart=0.0, Output_TRes=1, Output_TDelta=1.0, Int_TimeDirection=1,
                              FTLE_Compute=1, FTLE_MeshBounds_XMin=0.0, FTLE_MeshBounds_XMax=2.0,
                              FTLE_MeshBounds_YMin=0.0, FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_XRes=128,
                              FTLE_MeshBounds_YRes=64, FTLE_MeshBounds_ZRes=1, FTLE_IntTLength=10.0,
                              FTLE_OutFilePrefix="dgftle", FTLE_ICFile="dgftle_ic.bin")),
    # backward-time FTLE, 3 slides, 64x32 grid
    "dg_ftle_back": dict(gen=gen_dg, outputs=["dgb.0.bin", "dgb.1.bin", "dgb.2.bin", "dgb_noT.0.bin",
                                              "dgb_noT.1.bin", "dgb_noT.2.bin", "dgb_Cartesian.bin"],
                         deck=dict(DG, Output_TStart=10.0, Output_TRes=3, Output_TDelta=1.0, Int_TimeDirection=-1,
                                   FTLE_Compute=1, FTLE_MeshBounds_XMin=0.0, FTLE_MeshBounds_XMax=2.0,
                                   FTLE_MeshBounds_YMin=0.0, FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_XRes=64,
                                   FTLE_MeshBounds_YRes=32, FTLE_MeshBounds_ZRes=1, FTLE_IntTLength=4.0,
                                   FTLE_OutFilePrefix="dgb", FTLE_ICFile="dgb_ic.bin"), keep=["dgb_ic.bin"]),
    # 2-D Cartesian tracers from a generated 64x32 grid, 6 outputs.  The reference
    # leaves LeftDomainTime of generated Cartesian tracers uninitialised
    # (tracers.c:1340-1532); 2048 records (160 KB) come from a fresh zeroed
    # allocation, which the restatement reproduces by writing 0.
    "dg_tracers": dict(gen=gen_dg, outputs=["dgtr.%d.bin" % k for k in range(6)],
                       deck=dict(DG, Output_TStart=0.0, Output_TRes=6, Output_TDelta=2.0, Int_TimeDirection=1,
                                 Trace_Compute=1, Trace_GenerateMesh=1, Trace_MeshBounds_XMin=0.0,
                                 Trace_MeshBounds_XMax=2.0, Trace_MeshBounds_YMin=0.0, Trace_MeshBounds_YMax=1.0,
                                 Trace_MeshBounds_XRes=64, Trace_MeshBounds_YRes=32, Trace_MeshBounds_ZRes=1,
                                 Trace_IntTLength=10.0, Trace_OutFilePrefix="dgtr")),
    # 3-D unstructured tracers: 1000 seeds, one period, 11 outputs, exits
    "pipe_tracers": dict(gen=gen_pipe, outputs=["ptr.%d.bin" % k for k in range(11)], ic="ptr.IC",
                         deck=dict(PIPE, Output_TStart=0.0, Output_TRes=11, Output_TDelta=0.1, Int_TimeDirection=1,
                                   Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="seeds.bin",
                                   Trace_InFileFormat=4, Trace_IntTLength=1.0, Trace_OutFilePrefix="ptr")),
    # point location chain on special points (faces, edges, vertices, outside), no integration
    "pipe_locate": dict(gen=gen_pipe, outputs=[], ic="ploc.IC",
                        deck=dict(PIPE, Output_TStart=0.0, Output_TRes=1, Output_TDelta=0.1, Int_TimeDirection=1,
                                  Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="probes.txt",
                                  Trace_InFileFormat=1, Trace_IntTLength=0.0, Trace_OutFilePrefix="ploc")),
    # staggered release (tracers.c:293-796) from 21 points, unsteady flow (quadratic release-time solves);
    # one OpenMP thread so that the reference writes records in location order
    "pipe_staggered": dict(gen=gen_pipe_stagger, outputs=["pst.%d.bin" % k for k in range(9)],
                           deck=dict(STAGGER, Trace_OutFilePrefix="pst")),
    # staggered release in steady flow (linear release-time solves)
    "pipe_staggered_steady": dict(gen=gen_pipe_steady, outputs=["pss.%d.bin" % k for k in range(9)],
                                  deck=dict(STAGGER, Trace_OutFilePrefix="pss")),
    # activation potential (strain rate integrated along the paths): unstructured list seeds, one release
    "pipe_tracers_ap": dict(gen=gen_pipe_ap, outputs=["pap.%d.bin" % k for k in range(6)] +
                            ["pap_AP.0.bin", "pap_coordinates.bin"], ic="pap.IC",
                            deck=dict(PIPE, Output_TStart=0.0, Output_TRes=6, Output_TDelta=0.04, Int_TimeDirection=1,
                                      Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="seeds.bin",
                                      Trace_InFileFormat=4, Trace_IntTLength=0.15, Trace_APCompute=1,
                                      Trace_OutFilePrefix="pap")),
    # two releases: the reference's OutputAP reads each release file into a fresh allocation it then drops
    # (tracers.c:1649-1656, strainrate.c:372) and writes the stale array, so its _AP.1 repeats release 0;
    # the records with their Scalar in the position files are compared, and _AP.1 is checked against them
    "pipe_tracers_ap2": dict(gen=gen_pipe_ap, outputs=["pap2.%d.bin" % k for k in range(6)] +
                             ["pap2_AP.0.bin", "pap2_AP.1.bin", "pap2_coordinates.bin"], ic="pap2.IC",
                             skip=["pap2_AP.1.bin"],
                             deck=dict(PIPE, Output_TStart=0.0, Output_TRes=6, Output_TDelta=0.04,
                                       Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=0,
                                       Trace_InFile="seeds.bin", Trace_InFileFormat=4, Trace_IntTLength=0.15,
                                       Trace_NumLaunchTimes=2, Trace_LaunchTimeSpacing=0.04, Trace_APCompute=1,
                                       Trace_OutFilePrefix="pap2")),
    # exposure times per element (CET, MET, encounters) with 5 substeps per step; one OpenMP thread so that the
    # reference adds each element's times in one order
    "pipe_tracers_cet": dict(gen=gen_pipe, outputs=["pct.%d.bin" % k for k in range(3)] +
                             ["pcet_CET.0.bin", "pcet_MET.0.bin", "pcet_ENC.0.bin"], ic="pct.IC",
                             close=["pcet_CET.0.bin", "pcet_MET.0.bin"],
                             deck=dict(PIPE, Num_threads=1, Output_TStart=0.0, Output_TRes=3, Output_TDelta=0.05,
                                       Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=0,
                                       Trace_InFile="seeds.bin", Trace_InFileFormat=4, Trace_IntTLength=0.1,
                                       Trace_CETCompute=1, Trace_CETSubsteps=5, Trace_RTOutFilePrefix="pcet",
                                       Trace_OutFilePrefix="pct")),
    # residence times per seed (the location probes: points near the outlet leave, points outside stay -1)
    "pipe_tracers_rt": dict(gen=gen_pipe, outputs=["prtt.%d.bin" % k for k in range(4)] +
                            ["prt.0.bin", "prt_coordinates.bin"], ic="prtt.IC",
                            deck=dict(PIPE, Output_TStart=0.0, Output_TRes=4, Output_TDelta=0.1, Int_TimeDirection=1,
                                      Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="probes.txt",
                                      Trace_InFileFormat=1, Trace_IntTLength=0.3, Trace_RTCompute=1,
                                      Trace_RTOutFilePrefix="prt", Trace_OutFilePrefix="prtt")),
    # activation potential on the 2-D Cartesian grid (per-cell strain rate), generated seeds (2048 records: the
    # reference leaves LeftDomainTime of generated seeds unset, see dg_tracers)
    "dg_tracers_ap": dict(gen=gen_dg_ap, outputs=["dgap.%d.bin" % k for k in range(6)] +
                          ["dgap_AP.0.bin", "dgap_Cartesian.bin"],
                          deck=dict(DG, Output_TStart=0.0, Output_TRes=6, Output_TDelta=2.0, Int_TimeDirection=1,
                                    Trace_Compute=1, Trace_GenerateMesh=1, Trace_MeshBounds_XMin=0.0,
                                    Trace_MeshBounds_XMax=2.0, Trace_MeshBounds_YMin=0.0, Trace_MeshBounds_YMax=1.0,
                                    Trace_MeshBounds_XRes=64, Trace_MeshBounds_YRes=32, Trace_MeshBounds_ZRes=1,
                                    Trace_IntTLength=10.0, Trace_APCompute=1, Trace_OutFilePrefix="dgap")),
    # velocity output on a generated grid whose points are the mesh's voxel corners (every interior point on
    # element faces: the grid chain's guesses decide the elements), 5 output times
    "pipe_velout_grid": dict(gen=gen_pipe, outputs=["pvg.%d.bin" % k for k in range(5)] + ["pvg_Cartesian.bin"],
                             deck=dict(PIPE, **VELOUT, Output_TStart=0.0, Output_TRes=5, Output_TDelta=0.05,
                                       Int_TimeDirection=1, VelOut_GenerateMesh=1, VelOut_FilePrefix="pvg",
                                       VelOut_MeshBounds_XRes=13, VelOut_MeshBounds_YRes=13,
                                       VelOut_MeshBounds_ZRes=41)),
    # velocity output at listed points (ASCII list), with tracers in the same run
    "pipe_velout_list": dict(gen=gen_pipe_velout, outputs=["pvl.%d.bin" % k for k in range(4)] +
                             ["pvl_coordinates.bin"] + ["pvlt.%d.bin" % k for k in range(4)],
                             ic="pvlt.IC",
                             deck=dict(PIPE, **VELOUT, Output_TStart=0.0, Output_TRes=4, Output_TDelta=0.03,
                                       Int_TimeDirection=1, VelOut_GenerateMesh=0, VelOut_InFile="velpts.txt",
                                       VelOut_InFileFormat=1, VelOut_FilePrefix="pvl", Trace_Compute=1,
                                       Trace_GenerateMesh=0, Trace_InFile="seeds.bin", Trace_InFileFormat=4,
                                       Trace_IntTLength=0.09, Trace_OutFilePrefix="pvlt")),
    # 2-D Cartesian velocity output at listed points (some outside the data), backward output times
    "dg_velout_list": dict(gen=gen_dg_velout, outputs=["dgv.%d.bin" % k for k in range(6)] + ["dgv_coordinates.bin"],
                           deck=dict(DG, **VELOUT, Output_TStart=10.0, Output_TRes=6, Output_TDelta=1.5,
                                     Int_TimeDirection=-1, VelOut_GenerateMesh=0, VelOut_InFile="velpts.txt",
                                     VelOut_InFileFormat=1, VelOut_FilePrefix="dgv")),
    # Euler (integration.c:222-311) with extrapolation of points that leave (Advect, integration.c:22-45) and
    # Trace_AlwaysOutput, so that the extrapolated positions are written; 64x32 generated seeds (see dg_tracers)
    "dg_tracers_euler": dict(gen=gen_dg, outputs=["dge.%d.bin" % k for k in range(6)],
                             deck=dict(DG, Int_Type=0, Int_Extrapolate=1, Output_TStart=0.0, Output_TRes=6,
                                       Output_TDelta=2.0, Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=1,
                                       Trace_MeshBounds_XMin=0.0, Trace_MeshBounds_XMax=2.0,
                                       Trace_MeshBounds_YMin=0.0, Trace_MeshBounds_YMax=1.0,
                                       Trace_MeshBounds_XRes=64, Trace_MeshBounds_YRes=32, Trace_MeshBounds_ZRes=1,
                                       Trace_IntTLength=10.0, Trace_AlwaysOutput=1, Trace_OutFilePrefix="dge")),
    # Euler on the unstructured pipe: walk after each step, exit at the outlet
    "pipe_tracers_euler": dict(gen=gen_pipe, outputs=["pte.%d.bin" % k for k in range(11)], ic="pte.IC",
                               deck=dict(PIPE, Int_Type=0, Output_TStart=0.0, Output_TRes=11, Output_TDelta=0.1,
                                         Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=0,
                                         Trace_InFile="seeds.bin", Trace_InFileFormat=4, Trace_IntTLength=1.0,
                                         Trace_OutFilePrefix="pte")),
    # RK4 with extrapolation: velocity kept up to date after each step (integration.c:420-425), points that
    # leave move on with their last velocity, all written (Trace_AlwaysOutput); the location probes, whose
    # points near the outlet leave
    "pipe_tracers_extrap": dict(gen=gen_pipe, outputs=["ptx.%d.bin" % k for k in range(6)], ic="ptx.IC",
                                deck=dict(PIPE, Int_Extrapolate=1, Output_TStart=0.0, Output_TRes=6,
                                          Output_TDelta=0.1, Int_TimeDirection=1, Trace_Compute=1,
                                          Trace_GenerateMesh=0, Trace_InFile="probes.txt", Trace_InFileFormat=1,
                                          Trace_IntTLength=0.5, Trace_AlwaysOutput=1, Trace_OutFilePrefix="ptx")),
    # generated tracer grid on the pipe (tracers.c:1298-1560): x, y on voxel corners two voxels apart, z one voxel
    # apart from 0.5 below to 0.5 above the pipe, so inside points are mesh vertices and the end planes lie outside
    # the bounding box; Trace_AlwaysOutput writes those too (the reference leaves their LeftDomainTime unset:
    # 2205 records, 176 KB, come from fresh zeroed memory)
    "pipe_tracers_grid": dict(gen=gen_pipe, outputs=["ptg.%d.bin" % k for k in range(3)], ic="ptg.IC",
                              deck=dict(PIPE, Output_TStart=0.0, Output_TRes=3, Output_TDelta=0.05,
                                        Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=1,
                                        Trace_MeshBounds_XMin=-1.0, Trace_MeshBounds_XMax=1.0,
                                        Trace_MeshBounds_YMin=-1.0, Trace_MeshBounds_YMax=1.0,
                                        Trace_MeshBounds_ZMin=-0.5, Trace_MeshBounds_ZMax=10.5,
                                        Trace_MeshBounds_XRes=7, Trace_MeshBounds_YRes=7, Trace_MeshBounds_ZRes=45,
                                        Trace_IntTLength=0.1, Trace_AlwaysOutput=1, Trace_OutFilePrefix="ptg")),
    # backward-time tracers on the double gyre: two releases 1.0 apart from t=10 back to t=0 (tracers.c:1287-1291)
    "dg_tracers_back": dict(gen=gen_dg, outputs=["dgtb.%d.bin" % k for k in range(5)],  # no release reaches t=0
                            deck=dict(DG, Output_TStart=10.0, Output_TRes=6, Output_TDelta=2.0, Int_TimeDirection=-1,
                                      Trace_Compute=1, Trace_GenerateMesh=1, Trace_MeshBounds_XMin=0.0,
                                      Trace_MeshBounds_XMax=2.0, Trace_MeshBounds_YMin=0.0,
                                      Trace_MeshBounds_YMax=1.0, Trace_MeshBounds_XRes=64, Trace_MeshBounds_YRes=32,
                                      Trace_MeshBounds_ZRes=1, Trace_NumLaunchTimes=2, Trace_LaunchTimeSpacing=1.0,
                                      Trace_IntTLength=8.0, Trace_OutFilePrefix="dgtb")),
    # backward-time tracers on the pipe from t=1 (the periodic data's last frame wraps to file 0), exits
    # through the inlet
    "pipe_tracers_back": dict(gen=gen_pipe, outputs=["ptb.%d.bin" % k for k in range(6)], ic="ptb.IC",
                              deck=dict(PIPE, Output_TStart=1.0, Output_TRes=6, Output_TDelta=0.1,
                                        Int_TimeDirection=-1, Trace_Compute=1, Trace_GenerateMesh=0,
                                        Trace_InFile="seeds.bin", Trace_InFileFormat=4, Trace_IntTLength=0.5,
                                        Trace_OutFilePrefix="ptb")),
    # tracers over 1.6 periods of the periodic pipe data: frames past t=1 read the files again (velocity.c:184-190)
    "pipe_tracers_periods": dict(gen=gen_pipe, outputs=["ptp.%d.bin" % k for k in range(5)], ic="ptp.IC",
                                 deck=dict(PIPE, Output_TStart=0.2, Output_TRes=5, Output_TDelta=0.35,
                                           Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=0,
                                           Trace_InFile="seeds.bin", Trace_InFileFormat=4, Trace_IntTLength=1.4,
                                           Trace_OutFilePrefix="ptp")),
    # inward flow on the walls (Int_NormalFlow, velocity.c:213-228): wall-node velocities of every slice replaced
    # by 0.2 x the inward normal, forward and backward in time
    "pipe_tracers_normal": dict(gen=gen_pipe_normal, outputs=["ptn.%d.bin" % k for k in range(6)], ic="ptn.IC",
                                deck=dict(PIPE, Int_NormalFlow=1, Int_NormalFlowScaling=0.2, Output_TStart=0.0,
                                          Output_TRes=6, Output_TDelta=0.05, Int_TimeDirection=1, Trace_Compute=1,
                                          Trace_GenerateMesh=0, Trace_InFile="seeds.bin", Trace_InFileFormat=4,
                                          Trace_IntTLength=0.25, Trace_OutFilePrefix="ptn")),
    "pipe_tracers_normal_back": dict(gen=gen_pipe_normal, outputs=["ptnb.%d.bin" % k for k in range(6)],
                                     ic="ptnb.IC",
                                     deck=dict(PIPE, Int_NormalFlow=1, Int_NormalFlowScaling=0.2, Output_TStart=0.5,
                                               Output_TRes=6, Output_TDelta=0.05, Int_TimeDirection=-1,
                                               Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="seeds.bin",
                                               Trace_InFileFormat=4, Trace_IntTLength=0.25,
                                               Trace_OutFilePrefix="ptnb")),
    # three releases, each from its own seed file and .IC.<release> (Trace_MultipleInFiles)
    "pipe_tracers_multi": dict(gen=gen_pipe_multi, outputs=["ptm.%d.bin" % k for k in range(6)], ic="ptm.IC.0",
                               keep=["ptm.IC.1", "ptm.IC.2"],
                               deck=dict(PIPE, Output_TStart=0.0, Output_TRes=6, Output_TDelta=0.05,
                                         Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=0,
                                         Trace_InFile="mseeds", Trace_InFileFormat=4, Trace_MultipleInFiles=1,
                                         Trace_NumLaunchTimes=3, Trace_LaunchTimeSpacing=0.05,
                                         Trace_IntTLength=0.15, Trace_OutFilePrefix="ptm")),
    # FTLE written every 4 frames during the integration (FTLE_ComputeVariation, ftle.c:188-191, 1253-1325): files
    # numbered in output order, integration time so far, then the full-length FTLE
    "dg_ftle_var": dict(gen=gen_dg, outputs=["dgv%s.%d.bin" % (x, k) for x in ("", "_noT") for k in range(4)] +
                        ["dgv_Cartesian.bin"],
                        deck=dict(DG, Output_TStart=0.0, Output_TRes=1, Output_TDelta=1.0, Int_TimeDirection=1,
                                  FTLE_Compute=1, FTLE_MeshBounds_XMin=0.0, FTLE_MeshBounds_XMax=2.0,
                                  FTLE_MeshBounds_YMin=0.0, FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_XRes=64,
                                  FTLE_MeshBounds_YRes=32, FTLE_MeshBounds_ZRes=1, FTLE_IntTLength=6.0,
                                  FTLE_ComputeVariation=1, FTLE_VariationOutFreq=4, FTLE_OutFilePrefix="dgv",
                                  FTLE_ICFile="dgv_ic.bin")),
    # Runge-Kutta-Fehlberg (Int_Type=2, integration.c:464-699): adaptive steps between Int_MinTimeStep and
    # Int_MaxTimeStep at Int_Accuracy, on the double gyre (generated seeds) and the pipe (exits), and FTLE
    "dg_tracers_rkf": dict(gen=gen_dg, outputs=["dgr.%d.bin" % k for k in range(6)],
                           deck=dict(DG, Int_Type=2, Int_Accuracy=1e-6, Int_MinTimeStep=1e-4, Int_MaxTimeStep=0.05,
                                     Output_TStart=0.0, Output_TRes=6, Output_TDelta=2.0, Int_TimeDirection=1,
                                     Trace_Compute=1, Trace_GenerateMesh=1, Trace_MeshBounds_XMin=0.0,
                                     Trace_MeshBounds_XMax=2.0, Trace_MeshBounds_YMin=0.0,
                                     Trace_MeshBounds_YMax=1.0, Trace_MeshBounds_XRes=64, Trace_MeshBounds_YRes=32,
                                     Trace_MeshBounds_ZRes=1, Trace_IntTLength=10.0, Trace_OutFilePrefix="dgr")),
    "pipe_tracers_rkf": dict(gen=gen_pipe, outputs=["ptr2.%d.bin" % k for k in range(6)], ic="ptr2.IC",
                             deck=dict(PIPE, Int_Type=2, Int_Accuracy=1e-6, Int_MinTimeStep=1e-5, Int_MaxTimeStep=0.01,
                                       Output_TStart=0.0, Output_TRes=6, Output_TDelta=0.2, Int_TimeDirection=1,
                                       Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="probes.txt",
                                       Trace_InFileFormat=1, Trace_IntTLength=1.0, Trace_OutFilePrefix="ptr2")),
    # RKF with the staggered release, the activation potential and exposure times (integration.c:534-539, 650-677)
    # staggered release with normal flow: release times from the velocity as read, integration with the wall
    # velocities (main.c:55-67, tracers.c:488-495); release points near the wall
    "pipe_staggered_normal": dict(gen=gen_pipe_stagger_normal, outputs=["psn.%d.bin" % k for k in range(9)],
                                  deck=dict(STAGGER, Int_NormalFlow=1, Int_NormalFlowScaling=0.2,
                                            Trace_OutFilePrefix="psn")),
    "pipe_staggered_rkf": dict(gen=gen_pipe_stagger, outputs=["psr.%d.bin" % k for k in range(9)],
                               deck=dict(STAGGER, Int_Type=2, Int_Accuracy=1e-6, Int_MinTimeStep=1e-5,
                                         Int_MaxTimeStep=0.01, Trace_OutFilePrefix="psr")),
    "pipe_tracers_rkf_ap": dict(gen=gen_pipe_ap, outputs=["pra.%d.bin" % k for k in range(6)] +
                                ["pra_AP.0.bin", "pra_coordinates.bin"], ic="pra.IC",
                                deck=dict(PIPE, Int_Type=2, Int_Accuracy=1e-6, Int_MinTimeStep=1e-5,
                                          Int_MaxTimeStep=0.01, Output_TStart=0.0, Output_TRes=6, Output_TDelta=0.04,
                                          Int_TimeDirection=1, Trace_Compute=1, Trace_GenerateMesh=0,
                                          Trace_InFile="seeds.bin", Trace_InFileFormat=4, Trace_IntTLength=0.15,
                                          Trace_APCompute=1, Trace_OutFilePrefix="pra")),
    "pipe_tracers_rkf_cet": dict(gen=gen_pipe, outputs=["prc.%d.bin" % k for k in range(3)] +
                                 ["prcet_CET.0.bin", "prcet_MET.0.bin", "prcet_ENC.0.bin"], ic="prc.IC",
                                 close=["prcet_CET.0.bin", "prcet_MET.0.bin"],
                                 deck=dict(PIPE, Num_threads=1, Int_Type=2, Int_Accuracy=1e-6, Int_MinTimeStep=1e-5,
                                           Int_MaxTimeStep=0.01, Output_TStart=0.0, Output_TRes=3,
                                           Output_TDelta=0.05, Int_TimeDirection=1, Trace_Compute=1,
                                           Trace_GenerateMesh=0, Trace_InFile="seeds.bin", Trace_InFileFormat=4,
                                           Trace_IntTLength=0.1, Trace_CETCompute=1, Trace_CETSubsteps=5,
                                           Trace_RTOutFilePrefix="prcet", Trace_OutFilePrefix="prc")),
    "dg_ftle_rkf": dict(gen=gen_dg, outputs=["dgfr.0.bin", "dgfr_noT.0.bin", "dgfr_Cartesian.bin"],
                        deck=dict(DG, Int_Type=2, Int_Accuracy=1e-6, Int_MinTimeStep=1e-4, Int_MaxTimeStep=0.05,
                                  Output_TStart=0.0, Output_TRes=1, Output_TDelta=1.0, Int_TimeDirection=1,
                                  FTLE_Compute=1, FTLE_MeshBounds_XMin=0.0, FTLE_MeshBounds_XMax=2.0,
                                  FTLE_MeshBounds_YMin=0.0, FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_XRes=64,
                                  FTLE_MeshBounds_YRes=32, FTLE_MeshBounds_ZRes=1, FTLE_IntTLength=4.0,
                                  FTLE_OutFilePrefix="dgfr", FTLE_ICFile="dgfr_ic.bin")),
    # 2-D Cartesian FTLE with Euler, 64x32 grid, T=4 (early exits re-advected with Euler)
    "dg_ftle_euler": dict(gen=gen_dg, outputs=["dgfe.0.bin", "dgfe_noT.0.bin", "dgfe_Cartesian.bin"],
                          deck=dict(DG, Int_Type=0, Output_TStart=0.0, Output_TRes=1, Output_TDelta=1.0,
                                    Int_TimeDirection=1, FTLE_Compute=1, FTLE_MeshBounds_XMin=0.0,
                                    FTLE_MeshBounds_XMax=2.0, FTLE_MeshBounds_YMin=0.0, FTLE_MeshBounds_YMax=1.0,
                                    FTLE_MeshBounds_XRes=64, FTLE_MeshBounds_YRes=32, FTLE_MeshBounds_ZRes=1,
                                    FTLE_IntTLength=4.0, FTLE_OutFilePrefix="dgfe", FTLE_ICFile="dgfe_ic.bin")),
    # 3-D unstructured FTLE with extrapolation: points that leave move on with their last velocity and take part
    # in the gradient at the end of the slide (no early FTLE, ftle.c:113-170)
    "pipe_ftle_extrap": dict(gen=gen_pipe, outputs=["pfx.0.bin", "pfx.1.bin", "pfx_noT.0.bin", "pfx_noT.1.bin",
                                                    "pfx_Cartesian.bin"],
                             deck=dict(PIPE, Int_Extrapolate=1, Output_TStart=0.0, Output_TRes=2, Output_TDelta=0.1,
                                       Int_TimeDirection=1, FTLE_Compute=1, FTLE_MeshBounds_XMin=-1.0,
                                       FTLE_MeshBounds_XMax=1.0, FTLE_MeshBounds_YMin=-1.0, FTLE_MeshBounds_YMax=1.0,
                                       FTLE_MeshBounds_ZMin=0.0, FTLE_MeshBounds_ZMax=10.0, FTLE_MeshBounds_XRes=12,
                                       FTLE_MeshBounds_YRes=12, FTLE_MeshBounds_ZRes=30, FTLE_IntTLength=0.5,
                                       FTLE_OutFilePrefix="pfx", FTLE_ICFile="pfx_ic.bin")),
    # 3-D Cartesian tracers (trilinear interpolation, velocity.c:330-456) from a generated 16x16x10 grid
    # (2560 records: see dg_tracers for the reference's unset LeftDomainTime), exits through z=1
    "box_tracers": dict(gen=gen_box, outputs=["btr.%d.bin" % k for k in range(6)],
                        deck=dict(BOX, Output_TStart=0.0, Output_TRes=6, Output_TDelta=1.0, Int_TimeDirection=1,
                                  Trace_Compute=1, Trace_GenerateMesh=1, Trace_MeshBounds_XMin=0.05,
                                  Trace_MeshBounds_XMax=1.95, Trace_MeshBounds_YMin=0.05, Trace_MeshBounds_YMax=0.95,
                                  Trace_MeshBounds_ZMin=0.05, Trace_MeshBounds_ZMax=0.95, Trace_MeshBounds_XRes=16,
                                  Trace_MeshBounds_YRes=16, Trace_MeshBounds_ZRes=10, Trace_IntTLength=5.0,
                                  Trace_OutFilePrefix="btr")),
    # 3-D Cartesian FTLE on a 24x12x12 grid over the whole box (points on the walls), early exits, 2 slides
    "box_ftle": dict(gen=gen_box, outputs=["bft.0.bin", "bft.1.bin", "bft_noT.0.bin", "bft_noT.1.bin",
                                           "bft_Cartesian.bin"],
                     deck=dict(BOX, Output_TStart=0.0, Output_TRes=2, Output_TDelta=0.5, Int_TimeDirection=1,
                               FTLE_Compute=1, FTLE_MeshBounds_XMin=0.0, FTLE_MeshBounds_XMax=2.0,
                               FTLE_MeshBounds_YMin=0.0, FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_ZMin=0.0,
                               FTLE_MeshBounds_ZMax=1.0, FTLE_MeshBounds_XRes=24, FTLE_MeshBounds_YRes=12,
                               FTLE_MeshBounds_ZRes=12, FTLE_IntTLength=2.0, FTLE_OutFilePrefix="bft",
                               FTLE_ICFile="bft_ic.bin")),
    # 2-D unstructured tracers (triangles, velocity.c:570-620, mesh.c:566-605): random seeds, seeds on nodes and
    # edges, seeds outside; exits through the outlet
    # the reference's Stenosis2D data (real 2-D triangle mesh and velocity): location chain on random, vertex
    # and edge-midpoint probes; velocity output at listed points; tracers over two data frames with the
    # reference deck's time step (28.6 steps a frame: the last one clipped)
    "sten_locate": dict(gen=gen_sten_probes, outputs=[], ic="sloc.IC",
                        deck=dict(STEN, Output_TStart=1.414, Output_TRes=1, Output_TDelta=0.002, Int_TimeDirection=1,
                                  Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="probes.txt",
                                  Trace_InFileFormat=1, Trace_IntTLength=0.0, Trace_OutFilePrefix="sloc")),
    "sten_velout": dict(gen=gen_sten_velout, outputs=["svo.%d.bin" % k for k in range(4)] + ["svo_coordinates.bin"],
                        deck=dict(STEN, VelOut_Compute=1, VelOut_MeshBounds_XMin=-0.01715, VelOut_MeshBounds_XMax=0.01715,
                                  VelOut_MeshBounds_YMin=0.0, VelOut_MeshBounds_YMax=0.00275, Output_TStart=1.414,
                                  Output_TRes=4, Output_TDelta=0.0013, Int_TimeDirection=1, VelOut_GenerateMesh=0,
                                  VelOut_InFile="velpts.txt", VelOut_InFileFormat=1, VelOut_FilePrefix="svo")),
    "sten_tracers": dict(gen=gen_sten_seeds, outputs=["str.%d.bin" % k for k in range(5)], ic="str.IC",
                         deck=dict(STEN, Output_TStart=1.414, Output_TRes=5, Output_TDelta=0.001, Int_TimeDirection=1,
                                   Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="seeds.bin",
                                   Trace_InFileFormat=4, Trace_IntTLength=0.004, Trace_OutFilePrefix="str")),
    "chan_tracers": dict(gen=gen_chan, outputs=["ctr.%d.bin" % k for k in range(6)], ic="ctr.IC",
                         deck=dict(CHAN, Output_TStart=0.0, Output_TRes=6, Output_TDelta=0.2, Int_TimeDirection=1,
                                   Trace_Compute=1, Trace_GenerateMesh=0, Trace_InFile="seeds.bin",
                                   Trace_InFileFormat=4, Trace_IntTLength=1.0, Trace_OutFilePrefix="ctr")),
    # 2-D unstructured FTLE, 48x12 grid (points on the walls and on mesh edges), backward time, 2 slides
    "chan_ftle": dict(gen=gen_chan, outputs=["cft.0.bin", "cft.1.bin", "cft_noT.0.bin", "cft_noT.1.bin",
                                             "cft_Cartesian.bin"],
                      deck=dict(CHAN, Output_TStart=1.0, Output_TRes=2, Output_TDelta=0.1, Int_TimeDirection=-1,
                                FTLE_Compute=1, FTLE_MeshBounds_XMin=0.0, FTLE_MeshBounds_XMax=4.0,
                                FTLE_MeshBounds_YMin=0.0, FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_XRes=49,
                                FTLE_MeshBounds_YRes=13, FTLE_MeshBounds_ZRes=1, FTLE_IntTLength=0.5,
                                FTLE_OutFilePrefix="cft", FTLE_ICFile="cft_ic.bin")),
    # FTLE grid on the pipe's voxel corners (two voxels apart): every inside point is a mesh vertex held by many
    # elements, so each element id follows the location chain's guesses (ftle.c:456-601); kept: the grid records
    "pipe_ftle_vertices": dict(gen=gen_pipe, outputs=["pfv.0.bin", "pfv_noT.0.bin", "pfv_Cartesian.bin"],
                               keep=["pfv_ic.bin"],
                               deck=dict(PIPE, Output_TStart=0.0, Output_TRes=1, Output_TDelta=0.1,
                                         Int_TimeDirection=1, FTLE_Compute=1, FTLE_MeshBounds_XMin=-1.0,
                                         FTLE_MeshBounds_XMax=1.0, FTLE_MeshBounds_YMin=-1.0,
                                         FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_ZMin=0.0,
                                         FTLE_MeshBounds_ZMax=10.0, FTLE_MeshBounds_XRes=7, FTLE_MeshBounds_YRes=7,
                                         FTLE_MeshBounds_ZRes=21, FTLE_IntTLength=0.05, FTLE_OutFilePrefix="pfv",
                                         FTLE_ICFile="pfv_ic.bin")),
    # 3-D unstructured FTLE with outside masks and early exits, 2 slides
    "pipe_ftle": dict(gen=gen_pipe, outputs=["pft.0.bin", "pft.1.bin", "pft_noT.0.bin", "pft_noT.1.bin",
                                             "pft_Cartesian.bin"],
                      deck=dict(PIPE, Output_TStart=0.0, Output_TRes=2, Output_TDelta=0.1, Int_TimeDirection=1,
                                FTLE_Compute=1, FTLE_MeshBounds_XMin=-1.0, FTLE_MeshBounds_XMax=1.0,
                                FTLE_MeshBounds_YMin=-1.0, FTLE_MeshBounds_YMax=1.0, FTLE_MeshBounds_ZMin=0.0,
                                FTLE_MeshBounds_ZMax=10.0, FTLE_MeshBounds_XRes=12, FTLE_MeshBounds_YRes=12,
                                FTLE_MeshBounds_ZRes=30, FTLE_IntTLength=0.5, FTLE_OutFilePrefix="pft",
                                FTLE_ICFile="pft_ic.bin"), keep=["pft_ic.bin"]),
}

# loop choice of ompVC depends on the host processor count (parameters.c:398-421);
# the fixtures were written on an 8-processor container
GOLDEN_NPROC = 8


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def prepare(case, workdir, **over):
    """Generate inputs for `case` under workdir/data and return (deck_path,
    data_dir, out_dir); keyword arguments override deck entries."""
    c = CASES[case]
    data = os.path.join(workdir, "data")
    out = os.path.join(workdir, "out")
    os.makedirs(data, exist_ok=True)
    os.makedirs(out, exist_ok=True)
    c["gen"](data)
    dk = dict(c["deck"], Path_Data=data + "/", Path_Output=out + "/", **over)
    deck_path = os.path.join(workdir, case + ".in")
    deck.write_deck(deck_path, **dk)
    return deck_path, data, out


def input_digests(data):
    return {f: sha256(os.path.join(data, f)) for f in sorted(os.listdir(data))
            if not f.endswith(".IC") and ".IC." not in f and not f.endswith("_ic.bin") and "_slide" not in f and "_OUT" not in f
            and not f.startswith("pft.") and not f.startswith("dgftle.") and not f.startswith("dgb.")}


def fixture_dir(case):
    return os.path.join(HERE, case)


def load_manifest(case):
    with open(os.path.join(fixture_dir(case), "manifest.json")) as f:
        return json.load(f)
