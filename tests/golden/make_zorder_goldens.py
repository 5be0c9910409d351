"""Write golden fixtures for the mesh Z-order renumbering by running the
reference tool (cudaVC/ZOrderSort/main.cpp, built from its source by
oracle/ref.mk into oracle/_ref/zorder_sort) on small generated meshes.

Run in the authoring container only (needs /root/reference):
    make -f oracle/ref.mk && python tests/golden/make_zorder_goldens.py
Each fixture directory receives the reference's three renumbered mesh files
(coordinates, connectivity, adjacency; ompVC layouts with the leading count)
and manifest.json with the generator arguments and input digests.  The
velocity loop is not run (first > last): the reference writes the velocity
time stamp to a closed stream (main.cpp:282).
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import cases  # noqa: E402
from fvcpy import host  # noqa: E402

REF = os.path.join(cases.REPO, "oracle", "_ref", "zorder_sort")
MESH_FILES = ("coordinates", "connectivity", "adjacency")

# name -> (generator, arguments)
ZCASES = {
    "zorder_pipe": ("pipe", (6, 6, 10, 1.0, 10.0)),
    "zorder_ybif": ("ybif", (24, 9, 32)),
}


def make_mesh(name):
    gen, args = ZCASES[name]
    t = host.pipe(*args) if gen == "pipe" else host.ybif(*args)
    return host.tube_mesh(t)


def digest(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main(selected):
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    for name in selected or sorted(ZCASES):
        work = tempfile.mkdtemp(prefix="zgolden_")
        try:
            xyz, conn, adj = make_mesh(name)
            host.write_unstruct_mesh(work, "m", xyz, conn, adj)
            files = [os.path.join(work, "m_%s.bin" % k) for k in MESH_FILES]
            r = subprocess.run([REF] + files + [os.path.join(work, "m_vel"), "1", "0", "1"], cwd=work,
                               capture_output=True, text=True)
            if r.returncode:
                sys.exit("%s: reference failed\n%s\n%s" % (name, r.stdout[-2000:], r.stderr[-2000:]))
            fx = os.path.join(cases.HERE, name)
            if os.path.isdir(fx):
                shutil.rmtree(fx)
            os.makedirs(fx)
            for k, f in zip(MESH_FILES, files):
                shutil.copy(f + ".new", os.path.join(fx, "%s.bin.new" % k))
            man = dict(case=name, generator=ZCASES[name][0], args=list(ZCASES[name][1]),
                       inputs={k: digest(f) for k, f in zip(MESH_FILES, files)},
                       producer="cudaVC/ZOrderSort/main.cpp built by oracle/ref.mk (g++ -O2)")
            with open(os.path.join(fx, "manifest.json"), "w") as f:
                json.dump(man, f, indent=1, sort_keys=True)
            print(name, "->", xyz.shape[0], "nodes", conn.shape[0], "elements")
        finally:
            shutil.rmtree(work)


if __name__ == "__main__":
    main(sys.argv[1:])
